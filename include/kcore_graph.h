/*
 * kcore_graph.h — C ABI of libkcore_host.so: the edge-list -> CSR loader with the
 * reference's semantics, its serialized cache and core-number writers, and the
 * seeded synthetic graph generators used by the tests and the benchmark.
 *
 * Reference interfaces replaced (paths relative to the reference repository):
 *   kcg_load_edgelist      <- Graph::Graph / Graph::readFile        src/graph.cpp:58-124,142-154
 *   kcg_read_serialized    <- Graph::readSerialized                 src/graph.cpp:9-33
 *   kcg_write_serialized   <- Graph::writeSerialized                src/graph.cpp:35-54
 *   kcg_write_kcore_json   <- Graph::writeKCoreToDisk               src/graph.cpp:126-140
 *   kcg_graph_arrays       <- the public fields of class Graph      inc/graph.h:7-23
 * The generators have no reference counterpart (the reference ships no inputs);
 * they produce the same kind of input the reference's loader builds.
 *
 * Ownership: every kcg_graph* is owned by the library and released with
 * kcg_graph_free. Array pointers returned by kcg_graph_arrays stay valid until then.
 * No function calls exit(); each returns a status (0 = ok).
 */
#ifndef KCORE_GRAPH_H
#define KCORE_GRAPH_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct kcg_graph kcg_graph;

enum {
    KCG_OK = 0,
    KCG_ERR_IO = 1,       /* file missing / unwritable */
    KCG_ERR_RANGE = 2,    /* vertex id or arc count beyond uint32 */
    KCG_ERR_ARG = 3,      /* bad argument */
    KCG_ERR_FORMAT = 4    /* malformed serialized cache */
};

/* Load a whitespace-separated 'u v' edge list exactly as the reference loader does:
 * lines starting with an alphabetic char are skipped, self loops dropped,
 * V = 1 + max id seen, symmetrised, deduplicated, ascending adjacency.
 * nthreads <= 0 uses all host threads. */
int kcg_load_edgelist(const char* path, int nthreads, kcg_graph** out);

/* Serialized CSR cache in the reference's byte format ("V\nE\n" + degrees + offsets + neighbours). */
int kcg_read_serialized(const char* path, kcg_graph** out);
int kcg_write_serialized(const kcg_graph* g, const char* path);

/* Reference constructor behaviour: try <outdir>/serialized-<basename(path)> first, else parse
 * the text file and write that cache. *from_cache is set to 1 when the cache was used. */
int kcg_load_cached(const char* path, const char* outdir, int nthreads, kcg_graph** out, int* from_cache);

/* Binary cache (no reference counterpart; the text cache above stays for interop). File
 * <outdir>bin-<basename>-<16 hex digits of the source's content hash>.kcg: a 48-byte header
 * ("KCGBIN\0\1", version 1, V, arcs, source hash, payload hash, reserved) and the raw
 * little-endian offsets and neighbours. A cache whose header, hashes or offsets do not check out
 * is refused with KCG_ERR_FORMAT (kcg_load_cached_bin then parses the text and rewrites it). */
int kcg_hash_file(const char* path, uint64_t* out);
uint64_t kcg_hash_bytes(const void* p, uint64_t n, uint64_t seed);
int kcg_write_bincache(const kcg_graph* g, const char* path, uint64_t src_hash);
/* want_src_hash 0 accepts any source hash. */
int kcg_read_bincache(const char* path, uint64_t want_src_hash, kcg_graph** out);
int kcg_load_cached_bin(const char* path, const char* outdir, int nthreads, kcg_graph** out, int* from_cache);

/* Build a CSR from an explicit pair list (same rules as the text loader). */
int kcg_from_pairs(const uint32_t* src, const uint32_t* dst, uint64_t npairs, int nthreads, kcg_graph** out);

/* Wrap a caller-provided CSR (copied). */
int kcg_from_csr(uint32_t V, uint64_t arcs, const uint32_t* offsets, const uint32_t* neighbors, kcg_graph** out);

void kcg_graph_free(kcg_graph* g);
void kcg_graph_info(const kcg_graph* g, uint32_t* V, uint64_t* arcs, uint32_t* dmax);
/* Pointers into library memory: offsets[V+1], neighbors[arcs], degrees[V]. */
void kcg_graph_arrays(const kcg_graph* g, const uint32_t** offsets, const uint32_t** neighbors, const uint32_t** degrees);

/* JSON dictionary of non-zero core numbers, byte-identical to writeKCoreToDisk. */
int kcg_write_kcore_json(const char* path, uint32_t V, const uint32_t* core);
/* Raw little-endian uint32[V] core dump. */
int kcg_write_core_u32(const char* path, uint32_t V, const uint32_t* core);
/* Text core dump, one number per line, each followed by '\n': the form of the degree block of the
 * serialized cache (Graph::writeSerialized, src/graph.cpp:35-54). No reference counterpart. */
int kcg_write_core_text(const char* path, uint32_t V, const uint32_t* core);

/* ---- seeded synthetic generators (deterministic for a given seed, any thread count) ---- */

/* Erdos-Renyi G(n, m): the first m distinct non-loop pairs of a seeded uniform stream. */
int kcg_gen_er(uint32_t n, uint64_t m, uint64_t seed, int nthreads, kcg_graph** out);
/* Chung-Lu: weights ~ (i+1)^(-1/(gamma-1)) clipped to [wmin, wmax] and scaled to mean avg_deg;
 * round(n*avg_deg/2) endpoint pairs drawn proportional to weight; ids permuted by the seed. */
int kcg_gen_chung_lu(uint32_t n, double gamma, double wmin, double wmax, double avg_deg,
                     uint64_t seed, int nthreads, kcg_graph** out);
/* R-MAT: ndraws pairs on 2^scale ids with quadrant probabilities (a, b, c, 1-a-b-c);
 * ids permuted by the seed (Graph500 convention). */
int kcg_gen_rmat(uint32_t scale, uint64_t ndraws, double a, double b, double c,
                 uint64_t seed, int nthreads, kcg_graph** out);

/* Served-workload requests: request i of stream `seed` asks for m edges, m lognormal with the
 * given median and sigma (Box-Muller on counter-based draws), truncated to [mmin, mmax] by
 * redrawing, and a seeded coin flip picks the family (kind 0: R-MAT (0.57, 0.19, 0.19) on
 * 2^scale vertices, scale = ceil(log2(m/16)), with m draws; kind 1: Chung-Lu, exponent 2.1,
 * n = ceil(2m/12), average degree 2m/n, weights clipped to [1, min(2800, n/2)]). Deduplication
 * and self-loop removal leave somewhat fewer than m distinct edges. */
int kcg_request_shape(uint64_t seed, uint64_t i, double median_m, double sigma, uint64_t mmin, uint64_t mmax,
                      uint64_t* m_out, int* kind_out);
int kcg_gen_request(uint64_t seed, uint64_t i, double median_m, double sigma, uint64_t mmin, uint64_t mmax,
                    int nthreads, kcg_graph** out, int* kind_out);

/* Write the pair list a generator would feed the loader as a 'u v' text edge list
 * (one line per undirected edge u<v of the built CSR). */
int kcg_write_edgelist(const kcg_graph* g, const char* path);

/* splitmix64, exposed so the Python side derives per-graph seeds identically. */
uint64_t kcg_splitmix64(uint64_t x);

/* ---- batches: many graphs packed into one CSR (the layout of kc_batch_csr in kcore.h) ----
 * Graph g owns vertices [vtx_off[g], vtx_off[g+1]) of the packed vertex arrays (local ids
 * 0..V_g-1), arcs [arc_off[g], arc_off[g+1]) of the packed neighbour array (local ids) and the
 * V_g+1 local offsets starting at offsets[vtx_off[g] + g]. */
typedef struct kcg_batch kcg_batch;

/* G graphs; graph i has n_i = round(exp(U(ln nmin, ln nmax))) vertices and mean degree
 * U(dmin, dmax); even i are Erdos-Renyi G(n, round(n*avg/2)), odd i Chung-Lu with gamma
 * (weights clipped to [1, min(2800, n/2)]); every draw derives from (seed, i). */
int kcg_gen_batch(uint32_t G, uint32_t nmin, uint32_t nmax, double dmin, double dmax, double gamma,
                  uint64_t seed, int nthreads, kcg_batch** out);
/* Pack existing graphs (copied). */
int kcg_batch_from_graphs(const kcg_graph* const* graphs, uint32_t G, kcg_batch** out);
void kcg_batch_free(kcg_batch* b);
void kcg_batch_info(const kcg_batch* b, uint32_t* G, uint64_t* total_vertices, uint64_t* total_arcs);
void kcg_batch_arrays(const kcg_batch* b, const uint64_t** vtx_off, const uint64_t** arc_off,
                      const uint32_t** offsets, const uint32_t** neighbors, const uint32_t** degrees);
/* Copy graph g out of the batch as a stand-alone graph. */
int kcg_batch_graph(const kcg_batch* b, uint32_t g, kcg_graph** out);

#ifdef __cplusplus
}
#endif
#endif /* KCORE_GRAPH_H */
