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

int kcg_write_core_text(const char* path, uint32_t V, const uint32_t* core) {
    if (!path || (!core && V)) return KCG_ERR_ARG;
    return kcg::write_core_text(path, V, core);
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

// ---------------------------------------------------------------------------------------
// batches

#include <omp.h>

#include <cmath>
#include <cstring>
#include <vector>

struct kcg_batch {
    uint32_t G = 0;
    std::vector<uint64_t> vtx_off, arc_off;
    std::vector<uint32_t> offsets, neighbors, degrees;
};

static int pack_batch(const std::vector<const kcg::Csr*>& gs, kcg_batch** out) {
    kcg_batch* b = new (std::nothrow) kcg_batch;
    if (!b) return KCG_ERR_ARG;
    const uint32_t G = (uint32_t)gs.size();
    b->G = G;
    b->vtx_off.assign((size_t)G + 1, 0);
    b->arc_off.assign((size_t)G + 1, 0);
    for (uint32_t g = 0; g < G; ++g) {
        b->vtx_off[g + 1] = b->vtx_off[g] + gs[g]->V;
        b->arc_off[g + 1] = b->arc_off[g] + gs[g]->arcs;
    }
    b->offsets.resize(b->vtx_off[G] + G);
    b->neighbors.resize(b->arc_off[G]);
    b->degrees.resize(b->vtx_off[G]);
#pragma omp parallel for schedule(dynamic, 1)
    for (int64_t g = 0; g < (int64_t)G; ++g) {
        const kcg::Csr& c = *gs[g];
        std::memcpy(b->offsets.data() + b->vtx_off[g] + g, c.offsets.data(), ((size_t)c.V + 1) * 4);
        if (c.arcs) std::memcpy(b->neighbors.data() + b->arc_off[g], c.neighbors.data(), c.arcs * 4);
        if (c.V) std::memcpy(b->degrees.data() + b->vtx_off[g], c.degrees.data(), (size_t)c.V * 4);
    }
    *out = b;
    return KCG_OK;
}

extern "C" {

int kcg_gen_batch(uint32_t G, uint32_t nmin, uint32_t nmax, double dmin, double dmax, double gamma,
                  uint64_t seed, int nthreads, kcg_batch** out) {
    if (!out || G == 0 || nmin < 4 || nmax < nmin || dmin <= 0 || dmax < dmin) return KCG_ERR_ARG;
    std::vector<kcg::Csr> gs(G);
    std::vector<int> rc(G, 0);
    const int nt = nthreads > 0 ? nthreads : omp_get_max_threads();
    // graphs are generated in parallel, one thread each; each graph's draws depend only on (seed, i)
#pragma omp parallel for num_threads(nt) schedule(dynamic, 1)
    for (int64_t i = 0; i < (int64_t)G; ++i) {
        const uint64_t gs_seed = kcg::splitmix64(seed * 0x9E3779B97F4A7C15ull + (uint64_t)i);
        const uint64_t r1 = kcg::splitmix64(gs_seed ^ 0x51ED27A5u), r2 = kcg::splitmix64(r1);
        const double u1 = (double)(r1 >> 11) * (1.0 / 9007199254740992.0);
        const double u2 = (double)(r2 >> 11) * (1.0 / 9007199254740992.0);
        const uint32_t n = (uint32_t)std::llround(std::exp(std::log((double)nmin) +
                                                          u1 * (std::log((double)nmax) - std::log((double)nmin))));
        const double avg = dmin + u2 * (dmax - dmin);
        if (i % 2 == 0) {
            uint64_t m = (uint64_t)std::llround(n * avg / 2.0);
            const uint64_t cap = (uint64_t)n * (n - 1) / 4;
            if (m > cap) m = cap;
            rc[i] = kcg::gen_er(n, m, gs_seed, 1, gs[i]);
        } else {
            const double wmax = std::min(2800.0, n / 2.0);
            rc[i] = kcg::gen_chung_lu(n, gamma, 1.0, wmax, avg, gs_seed, 1, gs[i]);
        }
    }
    for (uint32_t i = 0; i < G; ++i)
        if (rc[i]) return rc[i];
    std::vector<const kcg::Csr*> ptrs(G);
    for (uint32_t i = 0; i < G; ++i) ptrs[i] = &gs[i];
    return pack_batch(ptrs, out);
}

uint64_t kcg_hash_bytes(const void* p, uint64_t n, uint64_t seed) { return kcg::hash_bytes(p, n, seed); }

int kcg_hash_file(const char* path, uint64_t* out) {
    if (!path || !out) return KCG_ERR_ARG;
    return kcg::hash_file(path, out);
}

int kcg_write_bincache(const kcg_graph* g, const char* path, uint64_t src_hash) {
    if (!g || !path) return KCG_ERR_ARG;
    return kcg::write_bincache(g->csr, path, src_hash);
}

int kcg_read_bincache(const char* path, uint64_t want_src_hash, kcg_graph** out) {
    if (!path || !out) return KCG_ERR_ARG;
    kcg::Csr c;
    const int st = kcg::read_bincache(path, c, want_src_hash);
    if (st != kcg::OK) return st;
    auto* g = new kcg_graph;
    g->csr = std::move(c);
    *out = g;
    return KCG_OK;
}

int kcg_load_cached_bin(const char* path, const char* outdir, int nthreads, kcg_graph** out, int* from_cache) {
    if (!path || !out) return KCG_ERR_ARG;
    kcg::Csr c;
    const int st = kcg::load_cached_bin(path, outdir ? outdir : "./output/", nthreads, c, from_cache);
    if (st != kcg::OK) return st;
    auto* g = new kcg_graph;
    g->csr = std::move(c);
    *out = g;
    return KCG_OK;
}

int kcg_request_shape(uint64_t seed, uint64_t i, double median_m, double sigma, uint64_t mmin, uint64_t mmax,
                      uint64_t* m_out, int* kind_out) {
    if (!m_out || !kind_out || median_m <= 0 || sigma < 0 || mmin == 0 || mmax < mmin) return KCG_ERR_ARG;
    const uint64_t key = kcg::splitmix64(seed * 0x9E3779B97F4A7C15ull + i);
    auto unif = [&](uint64_t j) {  // (0, 1]
        return ((double)(kcg::splitmix64(key + (j + 1) * 0x9E3779B97F4A7C15ull) >> 11) + 1.0) *
               (1.0 / 9007199254740992.0);
    };
    // lognormal by Box-Muller, truncated to [mmin, mmax] by redrawing (at most 64 tries, then clamped)
    double m = median_m;
    for (uint64_t t = 0; t < 64; ++t) {
        const double z = std::sqrt(-2.0 * std::log(unif(2 * t))) * std::cos(6.283185307179586 * unif(2 * t + 1));
        m = median_m * std::exp(sigma * z);
        if (m >= (double)mmin && m <= (double)mmax) break;
        m = std::min(std::max(m, (double)mmin), (double)mmax);
    }
    *m_out = (uint64_t)std::llround(m);
    *kind_out = (int)(kcg::splitmix64(key ^ 0xC0FFEEull) & 1);
    return KCG_OK;
}

int kcg_gen_request(uint64_t seed, uint64_t i, double median_m, double sigma, uint64_t mmin, uint64_t mmax,
                    int nthreads, kcg_graph** out, int* kind_out) {
    uint64_t m;
    int kind;
    int rc = kcg_request_shape(seed, i, median_m, sigma, mmin, mmax, &m, &kind);
    if (rc) return rc;
    if (!out) return KCG_ERR_ARG;
    const uint64_t gseed = kcg::splitmix64(seed * 0x9E3779B97F4A7C15ull + i) ^ 0x5EEDull;
    auto* g = new kcg_graph;
    if (kind == 0) {
        // R-MAT on 2^scale vertices, scale = ceil(log2(m / 16)) (edge factor 16), m draws,
        // (a, b, c) = (0.57, 0.19, 0.19)
        uint32_t scale = 6;
        while (scale < 30 && ((uint64_t)16 << scale) < m) ++scale;
        rc = kcg::gen_rmat(scale, m, 0.57, 0.19, 0.19, gseed, nthreads, g->csr);
    } else {
        // Chung-Lu, exponent 2.1, average degree 12 -> n = m/6, weights clipped to [1, min(2800, n/2)]
        const uint32_t n = (uint32_t)std::max<uint64_t>(64, (m + 3) / 6);
        const double avg = 2.0 * (double)m / n;
        rc = kcg::gen_chung_lu(n, 2.1, 1.0, std::min(2800.0, n / 2.0), avg, gseed, nthreads, g->csr);
    }
    if (rc) {
        delete g;
        return rc;
    }
    *out = g;
    if (kind_out) *kind_out = kind;
    return KCG_OK;
}

int kcg_batch_from_graphs(const kcg_graph* const* graphs, uint32_t G, kcg_batch** out) {
    if (!graphs || !out || G == 0) return KCG_ERR_ARG;
    std::vector<const kcg::Csr*> ptrs(G);
    for (uint32_t i = 0; i < G; ++i) {
        if (!graphs[i]) return KCG_ERR_ARG;
        ptrs[i] = &graphs[i]->csr;
    }
    return pack_batch(ptrs, out);
}

void kcg_batch_free(kcg_batch* b) { delete b; }

void kcg_batch_info(const kcg_batch* b, uint32_t* G, uint64_t* total_vertices, uint64_t* total_arcs) {
    if (G) *G = b->G;
    if (total_vertices) *total_vertices = b->vtx_off[b->G];
    if (total_arcs) *total_arcs = b->arc_off[b->G];
}

void kcg_batch_arrays(const kcg_batch* b, const uint64_t** vtx_off, const uint64_t** arc_off,
                      const uint32_t** offsets, const uint32_t** neighbors, const uint32_t** degrees) {
    if (vtx_off) *vtx_off = b->vtx_off.data();
    if (arc_off) *arc_off = b->arc_off.data();
    if (offsets) *offsets = b->offsets.data();
    if (neighbors) *neighbors = b->neighbors.data();
    if (degrees) *degrees = b->degrees.data();
}

int kcg_batch_graph(const kcg_batch* b, uint32_t g, kcg_graph** out) {
    if (!b || !out || g >= b->G) return KCG_ERR_ARG;
    const uint32_t V = (uint32_t)(b->vtx_off[g + 1] - b->vtx_off[g]);
    const uint64_t arcs = b->arc_off[g + 1] - b->arc_off[g];
    kcg::Csr c;
    c.V = V;
    c.arcs = arcs;
    const uint32_t* off = b->offsets.data() + b->vtx_off[g] + g;
    c.offsets.assign(off, off + (size_t)V + 1);
    c.neighbors.assign(b->neighbors.data() + b->arc_off[g], b->neighbors.data() + b->arc_off[g] + arcs);
    c.degrees.assign(b->degrees.data() + b->vtx_off[g], b->degrees.data() + b->vtx_off[g] + V);
    c.dmax = 0;
    for (uint32_t v = 0; v < V; ++v) c.dmax = std::max(c.dmax, c.degrees[v]);
    return wrap(std::move(c), out);
}

}  // extern "C"
