// graph_capi.cpp — extern "C" surface of libkcore_host.so (declared in include/kcore_graph.h).
#include <new>
#include <string>

#include "../../include/kcore_graph.h"
#include "graph.h"

struct kcg_graph {
    kcg::Csr csr;
};

static int wrap(kcg::Csr&& c, kcg_graph** out) {
    kcg_graph* g = new (std::nothrow) kcg_graph;
    if (!g) return KCG_ERR_ARG;
    g->csr = std::move(c);
    *out = g;
    return KCG_OK;
}

extern "C" {

int kcg_load_edgelist(const char* path, int nthreads, kcg_graph** out) {
    if (!path || !out) return KCG_ERR_ARG;
    kcg::Csr c;
    int st = kcg::load_edgelist(path, nthreads, c);
    return st ? st : wrap(std::move(c), out);
}

int kcg_read_serialized(const char* path, kcg_graph** out) {
    if (!path || !out) return KCG_ERR_ARG;
    kcg::Csr c;
    int st = kcg::read_serialized(path, c);
    return st ? st : wrap(std::move(c), out);
}

int kcg_write_serialized(const kcg_graph* g, const char* path) {
    if (!g || !path) return KCG_ERR_ARG;
    return kcg::write_serialized(g->csr, path);
}

int kcg_load_cached(const char* path, const char* outdir, int nthreads, kcg_graph** out,
                    int* from_cache) {
    if (!path || !outdir || !out) return KCG_ERR_ARG;
    std::string dir(outdir);
    if (!dir.empty() && dir.back() != '/') dir.push_back('/');
    std::string cache = dir + "serialized-" + kcg::basename_of(path);
    kcg::Csr c;
    if (kcg::read_serialized(cache, c) == kcg::OK) {
        if (from_cache) *from_cache = 1;
        return wrap(std::move(c), out);
    }
    if (from_cache) *from_cache = 0;
    int st = kcg::load_edgelist(path, nthreads, c);
    if (st) return st;
    kcg::write_serialized(c, cache);  // the reference writes it unconditionally; failure only warns
    return wrap(std::move(c), out);
}

int kcg_from_pairs(const uint32_t* src, const uint32_t* dst, uint64_t npairs, int nthreads,
                   kcg_graph** out) {
    if ((!src || !dst) && npairs) return KCG_ERR_ARG;
    if (!out) return KCG_ERR_ARG;
    kcg::Csr c;
    int st = kcg::build_csr(src, dst, npairs, nthreads, c);
    return st ? st : wrap(std::move(c), out);
}

int kcg_from_csr(uint32_t V, uint64_t arcs, const uint32_t* offsets, const uint32_t* neighbors,
                 kcg_graph** out) {
    if (!offsets || (!neighbors && arcs) || !out) return KCG_ERR_ARG;
    if (arcs > 0xffffffffull || offsets[0] != 0 || offsets[V] != arcs) return KCG_ERR_RANGE;
    kcg::Csr c;
    c.V = V;
    c.arcs = arcs;
    c.offsets.assign(offsets, offsets + (size_t)V + 1);
    c.neighbors.assign(neighbors, neighbors + arcs);
    c.degrees.resize(V);
    c.dmax = 0;
    for (uint32_t v = 0; v < V; ++v) {
        if (offsets[v + 1] < offsets[v]) return KCG_ERR_RANGE;
        c.degrees[v] = offsets[v + 1] - offsets[v];
        if (c.degrees[v] > c.dmax) c.dmax = c.degrees[v];
    }
    for (uint64_t j = 0; j < arcs; ++j)
        if (neighbors[j] >= V) return KCG_ERR_RANGE;
    return wrap(std::move(c), out);
}

void kcg_graph_free(kcg_graph* g) { delete g; }

void kcg_graph_info(const kcg_graph* g, uint32_t* V, uint64_t* arcs, uint32_t* dmax) {
    if (V) *V = g->csr.V;
    if (arcs) *arcs = g->csr.arcs;
    if (dmax) *dmax = g->csr.dmax;
}

void kcg_graph_arrays(const kcg_graph* g, const uint32_t** offsets, const uint32_t** neighbors,
                      const uint32_t** degrees) {
    if (offsets) *offsets = g->csr.offsets.data();
    if (neighbors) *neighbors = g->csr.neighbors.data();
    if (degrees) *degrees = g->csr.degrees.data();
}

int kcg_write_kcore_json(const char* path, uint32_t V, const uint32_t* core) {
    if (!path || (!core && V)) return KCG_ERR_ARG;
    return kcg::write_kcore_json(path, V, core);
}

int kcg_write_core_u32(const char* path, uint32_t V, const uint32_t* core) {
    if (!path || (!core && V)) return KCG_ERR_ARG;
    return kcg::write_core_u32(path, V, core);
}

int kcg_gen_er(uint32_t n, uint64_t m, uint64_t seed, int nthreads, kcg_graph** out) {
    if (!out) return KCG_ERR_ARG;
    kcg::Csr c;
    int st = kcg::gen_er(n, m, seed, nthreads, c);
    return st ? st : wrap(std::move(c), out);
}

int kcg_gen_chung_lu(uint32_t n, double gamma, double wmin, double wmax, double avg_deg,
                     uint64_t seed, int nthreads, kcg_graph** out) {
    if (!out) return KCG_ERR_ARG;
    kcg::Csr c;
    int st = kcg::gen_chung_lu(n, gamma, wmin, wmax, avg_deg, seed, nthreads, c);
    return st ? st : wrap(std::move(c), out);
}

int kcg_gen_rmat(uint32_t scale, uint64_t ndraws, double a, double b, double c, uint64_t seed,
                 int nthreads, kcg_graph** out) {
    if (!out) return KCG_ERR_ARG;
    kcg::Csr g;
    int st = kcg::gen_rmat(scale, ndraws, a, b, c, seed, nthreads, g);
    return st ? st : wrap(std::move(g), out);
}

int kcg_write_edgelist(const kcg_graph* g, const char* path) {
    if (!g || !path) return KCG_ERR_ARG;
    return kcg::write_edgelist(g->csr, path);
}

uint64_t kcg_splitmix64(uint64_t x) { return kcg::splitmix64(x); }

}  // extern "C"
