/*
 * kcore.h — C ABI of libkcore_hip.so, the MI355X (gfx950) k-core peeling engine.
 *
 * The reference exposes its GPU path as nine `int kcoreX(Graph&)` host functions
 * (src/ours.cc:117-160, src/ours-shared.cc:119, src/ours-prefetch.cc:153, src/ballot*.cc,
 * src/efficient*.cc), each doing cudaMalloc + H2D of the CSR (src/gpu_memory_allocation.cu:4-12),
 * the level loop, and returning elapsed ms, with core numbers left on the device
 * (get_results_from_gpu, src/gpu_memory_allocation.cu:14-16, never called). This ABI splits
 * that into a resident device graph and a peel call that returns core numbers:
 *
 *   kc_open / kc_close           <- (implicit CUDA context of the reference)
 *   kc_graph_upload              <- malloc_graph_gpu_memory   src/gpu_memory_allocation.cu:4-12
 *   kc_graph_free                <- free_graph_gpu_memory     src/gpu_memory_allocation.cu:18-22
 *   kc_peel                      <- kcore / kcoreSharedMem / kcorePrefetch (+ BC/EC clones)
 *                                   src/ours.cc:117-160 (variant chosen by kc_params.variant)
 *   kc_graph_download_core       <- get_results_from_gpu      src/gpu_memory_allocation.cu:14-16
 *   kc_graph_download_order      <- KCore glBuffer + cShellOffsets (peel order, k-shell starts)
 *                                   kcore-global-buffer/KCore.h:244-272
 *   kc_batch_*                   <- (no reference counterpart: batched peeling of many graphs)
 *   kc_serve_*                   <- (no reference counterpart: continuous-batching server)
 *   kc_status_str                <- chkerr's message print   inc/gpu_memory_allocation.h:9-16
 *
 * Conventions: the caller owns host buffers; the library owns device state; handles are
 * opaque; no function exits the process; every call returns a kc_status. A peel never
 * modifies the uploaded CSR or degrees (it peels a working copy), so repeated peels and
 * variant comparisons start from the same input. One kc_ctx per host thread.
 */
#ifndef KCORE_H
#define KCORE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int kc_status;
enum {
    KC_OK = 0,
    KC_ERR_ARG = 1,         /* invalid argument or graph shape */
    KC_ERR_HIP = 2,         /* a HIP runtime call failed */
    KC_ERR_OOM = 3,         /* device allocation failed */
    KC_ERR_CAPACITY = 4,    /* request beyond a documented maximum */
    KC_ERR_NO_PROGRESS = 5, /* level loop passed the maximum degree with vertices left */
    KC_ERR_DEVICE = 6,      /* a device-side wait exceeded its time limit */
    KC_ERR_NO_DEVICE = 7    /* no usable GPU */
};

typedef struct kc_ctx kc_ctx;
typedef struct kc_graph kc_graph;

/* Host CSR as the reference's Graph holds it (inc/graph.h:9-15, G_pointers inc/common.h:45-50):
 * offsets[V+1], neighbors[arcs] (ascending, duplicate-free, symmetric), degrees[V]
 * (may be NULL: taken from offsets). arcs must be < 2^32. */
typedef struct {
    uint32_t V;
    uint64_t arcs;
    const uint32_t* offsets;
    const uint32_t* neighbors;
    const uint32_t* degrees;
} kc_csr;

/* Peel variants. Every variant produces identical core numbers; they differ in how the
 * frontier is staged. The reference's labels map as: Ours -> WAVE, SM -> LDS, VP -> PREFETCH
 * (main.cu:60-88); its BC/EC compaction flavours are the wave64 ballot compaction used by all. */
enum {
    KC_VARIANT_AUTO = 0,
    KC_VARIANT_WAVE = 1,     /* one persistent kernel; each wave drains global frontier tickets */
    KC_VARIANT_LDS = 2,      /* + workgroup LDS frontier buffer for cascades, spilling to global */
    KC_VARIANT_PREFETCH = 3, /* + next-batch adjacency prefetch while the current batch runs */
    KC_VARIANT_STEPWISE = 4, /* reference-shaped: the host runs the level loop, launching a select
                                kernel and process kernels per level and reading the frontier size
                                back after each (src/ours.cc:117-160); a baseline, not a fast path */
    KC_VARIANT_ASYNC = 5     /* one persistent kernel; two grid barriers per level: the level's
                                cascade runs from a shared work queue with no barrier inside it,
                                and scans read a compacted list once few vertices are left */
};

typedef struct {
    int variant;              /* KC_VARIANT_* */
    uint32_t block;           /* threads per workgroup (256, 512 or 1024); 0 = tuned default */
    uint32_t lds_depth;       /* LDS frontier entries per workgroup (LDS variant); 0 = default */
    uint32_t heavy_threshold; /* degree above which a vertex's arcs are split into chunks shared
                                 by many waves; 0 = default, UINT32_MAX = never split */
    uint32_t flags;           /* KC_FLAG_* */
    uint32_t chunk_arcs;      /* arcs per chunk; 0 = default */
    uint32_t compact_ratio;   /* ASYNC: a scan compacts the alive list when unpeeled vertices *
                                 compact_ratio < entries it reads; 0 = default (4), UINT32_MAX =
                                 never compact */
    uint32_t poll_max;        /* ASYNC: longest idle poll sleep step, 0..3; 0 = default (2) */
} kc_params;

enum {
    KC_FLAG_NO_LEVEL_SKIP = 1u << 0, /* scan every level 0..kmax like the reference, even empty ones */
    KC_FLAG_TRACE = 1u << 8,         /* record a device timestamp at every grid barrier */
    KC_FLAG_GRAPH = 1u << 9,         /* kc_batch_peel: launch the whole peel (state reset, degree
                                        copy, persistent kernel) as one HIP graph, captured on first
                                        use for each (block, flags) and replayed afterwards */
    KC_FLAG_DIRECT_POLL = 1u << 10,  /* ASYNC: idle workgroups poll the shared queue counters
                                        directly instead of their XCD group's work hint */
    KC_FLAG_LEVEL_HANDOFF = 1u << 11 /* ASYNC: no grid barriers — each workgroup's scan is a work
                                        item, processing waits only for the level's scan count,
                                        and the thread that finds a level complete publishes the
                                        next one. Experimental: never chosen by the tuner */
};

typedef struct {
    double ms_peel;           /* device time from "CSR + degrees resident" to "cores + kmax resident" */
    double ms_h2d;            /* upload time of the graph (set by kc_graph_upload's caller stats) */
    double ms_d2h;            /* core download time when kc_peel copies to the host */
    double ms_kernel;         /* device time of the peeling kernel(s) alone, without state reset */
    uint32_t kmax;
    uint32_t levels;          /* levels whose degree==level scan ran */
    uint32_t subrounds;       /* processing sub-rounds (grid-synchronous steps) over all levels */
    uint64_t scanned_entries; /* degree entries read by the select scans */
    uint64_t bytes_alg;       /* algorithmic bytes: 4*scanned + 16*n + 20*m */
    uint64_t bytes_scan;      /* 4*scanned_entries */
    uint64_t bytes_edge;      /* 16*n + 20*m */
    uint32_t grid;            /* workgroups launched */
    uint32_t block;           /* threads per workgroup */
    uint32_t participants;    /* workgroups resident at start that ran the peel (<= grid) */
} kc_stats;

const char* kc_status_str(kc_status s);

kc_status kc_open(int device, kc_ctx** out);
void kc_close(kc_ctx* ctx);

/* Copies the CSR to the device, checks it there (offsets monotone from 0 to arcs, neighbour ids
 * < V: KC_ERR_ARG otherwise) and allocates the peel state for it. KC_ERR_CAPACITY when
 * V > 2^30 - 256 or arcs >= 2^32. */
kc_status kc_graph_upload(kc_ctx* ctx, const kc_csr* g, kc_graph** out);
void kc_graph_free(kc_graph* g);
void kc_graph_info(const kc_graph* g, uint32_t* V, uint64_t* arcs, uint32_t* dmax);

/* Peels the graph. core_out (host, V entries) and kmax_out / stats may be NULL; the core
 * numbers also stay on the device until the next peel of the same graph. */
kc_status kc_peel(kc_ctx* ctx, kc_graph* g, const kc_params* p, uint32_t* core_out,
                  uint32_t* kmax_out, kc_stats* stats);

/* Results of the last peel: core numbers (V entries), peel order (V entries: vertices in the
 * order they were peeled) and k-shell starts (shell_off[k] = first position in the order of a
 * vertex with core k, for k = 0..kmax+1; kmax+2 entries). */
kc_status kc_graph_download_core(kc_ctx* ctx, const kc_graph* g, uint32_t* core_out);
kc_status kc_graph_download_order(kc_ctx* ctx, const kc_graph* g, uint32_t* order_out,
                                  uint32_t* shell_off_out);
/* Device pointer of the core array (valid until kc_graph_free), for zero-copy consumers. */
const uint32_t* kc_graph_core_device(const kc_graph* g);

/* ---- batched peeling (no reference counterpart) ----
 * Many independent graphs packed into one CSR: graph g owns vertices [vtx_off[g], vtx_off[g+1])
 * of the packed vertex arrays (local ids 0..V_g-1), arcs [arc_off[g], arc_off[g+1]) of the packed
 * neighbour array (local ids), and V_g+1 local offsets starting at offsets[vtx_off[g] + g]
 * (offsets relative to arc_off[g]). One persistent kernel peels all of them, one workgroup per
 * graph at a time, largest graphs first, each graph with its own level counter. */
typedef struct {
    uint32_t G;
    uint64_t total_vertices;     /* vtx_off[G] */
    uint64_t total_arcs;         /* arc_off[G] */
    const uint64_t* vtx_off;     /* G+1 */
    const uint64_t* arc_off;     /* G+1 */
    const uint32_t* offsets;     /* total_vertices + G */
    const uint32_t* neighbors;   /* total_arcs */
    const uint32_t* degrees;     /* total_vertices; NULL = from offsets */
} kc_batch_csr;
typedef struct kc_batch kc_batch;

/* Copies the batch to the device and checks it there (offsets monotone, local ids < V_g). */
kc_status kc_batch_upload(kc_ctx* ctx, const kc_batch_csr* b, kc_batch** out);
void kc_batch_free(kc_batch* b);
/* core_out: total_vertices entries (packed like the input), kmax_out: G entries; both may be
 * NULL. stats: levels / subrounds summed over graphs, bytes_alg summed over graphs. Only
 * kc_params.block and .flags (KC_FLAG_NO_LEVEL_SKIP, KC_FLAG_GRAPH) apply. With KC_FLAG_GRAPH,
 * stats.ms_kernel equals stats.ms_peel (the graph is timed as a whole). */
kc_status kc_batch_peel(kc_ctx* ctx, kc_batch* b, const kc_params* p, uint32_t* core_out,
                        uint32_t* kmax_out, kc_stats* stats);
/* Per-graph peel order, packed like the vertex arrays (local ids). */
kc_status kc_batch_download_order(kc_ctx* ctx, const kc_batch* b, uint32_t* order_out);
void kc_batch_info(const kc_batch* b, uint32_t* G, uint64_t* total_vertices, uint64_t* total_arcs);

/* ---- autotuner (no reference counterpart: the reference hard-codes BLOCK_DIM, BUFF_SIZE and
 * its warp/block split in inc/common.h) ----
 * Times real peels of the resident graph (batch) and returns the fastest kc_params: for a graph,
 * coordinate descent over variant, block, heavy_threshold (the wave/block split), chunk_arcs and
 * lds_depth starting from the defaults; for a batch, block. Results must be identical for every
 * candidate: a single-graph candidate is adopted only if its core numbers equal those of the
 * default settings, and one that fails with a device-side status is skipped. cache_path (may be NULL) names a text file of tuned settings keyed by graph shape,
 * flags and device; a cached key returns at once with *trials = 0. */
typedef struct {
    uint32_t reps;          /* timed peels per candidate (median); 0 = 2 */
    uint32_t flags;         /* KC_FLAG_* every candidate uses */
    const char* cache_path;
} kc_tune_cfg;
kc_status kc_tune_peel(kc_ctx* ctx, kc_graph* g, const kc_tune_cfg* cfg, kc_params* best, double* best_ms,
                       uint32_t* trials);
kc_status kc_tune_batch(kc_ctx* ctx, kc_batch* b, const kc_tune_cfg* cfg, kc_params* best, double* best_ms,
                        uint32_t* trials);

/* ---- served peeling (no reference counterpart) ----
 * A server keeps one persistent kernel resident on the device. Each request is a graph as a host
 * CSR (offsets and neighbours; kc_csr.degrees is not used, degrees come from the offsets). It is
 * admitted, in arrival order, when its ring slot is free and the HBM arena has room for its
 * charge, max(8 m + 16 n bytes, its region) for n vertices and m undirected edges (the region:
 * offsets, neighbours and two vertex arrays, each rounded up to 256 bytes). Its CSR is copied
 * into the region on a copy stream and published in the ring; an idle workgroup of the kernel
 * takes it at once, checks it (offsets monotone from 0 to arcs, neighbour ids < V: a malformed
 * request completes with KC_ERR_ARG and is never peeled) and peels it alone — requests join and
 * leave the running kernel individually. When core_out is pinned host memory (kc_host_alloc)
 * the kernel writes the core numbers there itself; otherwise kc_serve_poll copies them down.
 * kc_serve_poll reports a request once its core numbers are in core_out and frees its region.
 * Requests of at least large_arcs arcs are peeled instead by the single-graph kernel over many
 * workgroups, on their own streams next to the resident kernel, with their whole single-graph
 * state (about 4.7 arcs + 28 n bytes + 1.3 MiB) placed in their own arena region: they are
 * charged max(8 m + 16 n, that state) and allocate no other device memory.
 * The CSR arrays and core_out must stay valid until the request's completion is returned.
 * Not thread-safe: one host thread drives a server.
 * The kernel never outlives the process: idle workgroups exit when the library's heartbeat
 * thread has been silent for lease_ms, and open servers are stopped at process exit.
 * While a serving kernel is resident, hipFree, hipHostFree and hipDeviceSynchronize called by
 * the process wait for it (they synchronise the whole device) and so do not return before the
 * server is closed; this library defers its own frees (kc_graph_free, kc_batch_free,
 * kc_host_free) until no serving kernel is resident. */
typedef struct kc_server kc_server;
typedef struct {
    uint64_t hbm_budget_bytes; /* arena size; 0 = 8 GiB */
    uint32_t ring_slots;       /* most requests admitted at once; 0 = 1024 */
    uint32_t block;            /* 256, 512 or 1024 threads per workgroup; 0 = 256 */
    uint32_t workgroups;       /* resident serving workgroups; 0 = 2 per CU (256) or 1 per CU */
    uint32_t flags;            /* KC_FLAG_NO_LEVEL_SKIP */
    uint32_t lease_ms;         /* heartbeat lease of the resident kernel; 0 = 5000 */
    uint32_t large_arcs;       /* requests with at least this many arcs go to a large-request lane
                                  (the single-graph kernel over many workgroups, next to the
                                  resident kernel) instead of one workgroup; 0 = 1M, UINT32_MAX =
                                  never */
    uint32_t large_lanes;      /* large requests peeled at once; 0 = 2 */
} kc_serve_cfg;
typedef struct {
    uint64_t req_id;
    kc_status status;          /* KC_OK; KC_ERR_ARG for a malformed CSR; KC_ERR_NO_PROGRESS
                                  when the level loop overran */
    uint32_t kmax, levels, rounds;
    double t_submit, t_admit, t_done; /* kc_serve_clock() seconds */
} kc_completion;
kc_status kc_serve_open(kc_ctx* ctx, const kc_serve_cfg* cfg, kc_server** out);
/* KC_ERR_CAPACITY: V >= 2^30, arcs >= 2^32 or a footprint above the whole arena. */
kc_status kc_serve_submit(kc_server* s, uint64_t req_id, const kc_csr* g, uint32_t* core_out);
/* Writes up to cap completions; *n_out = how many. Also admits waiting requests.
 * KC_ERR_DEVICE: the kernel has ended (its lease ran out); the server must be closed. */
kc_status kc_serve_poll(kc_server* s, kc_completion* out, uint32_t cap, uint32_t* n_out);
kc_status kc_serve_pending(const kc_server* s, uint32_t* queued, uint32_t* inflight, uint64_t* arena_used);
/* Stops the kernel after every admitted request has run; KC_ERR_DEVICE if it does not exit. */
kc_status kc_serve_close(kc_server* s);
double kc_serve_clock(void);
/* Pinned host memory (fast asynchronous copies; written directly by the serving kernel) for request
 * buffers, in 4 KiB multiples. kc_host_free takes only kc_host_alloc blocks (else KC_ERR_ARG,
 * also for a second free); a block freed while a serving kernel is resident is kept and handed
 * out again by kc_host_alloc, and released once no serving kernel is resident. Thread-safe. */
kc_status kc_host_alloc(uint64_t bytes, void** out);
kc_status kc_host_free(void* p);

/* Diagnostics (no reference counterpart). kc_selftest checks the wave64 DPP prefix sum the
 * kernels use against a host prefix sum on one wavefront (KC_OK when they agree).
 * kc_graph_download_trace returns the barrier timestamps of the last peel run with
 * KC_FLAG_TRACE: pairs (kind << 56 | subround << 32 | level, wall-clock ticks), kind 1 = after
 * select, 2 = after a sub-round, 3 = level end; ticks run at kc_wall_clock_khz kHz. */
kc_status kc_selftest(kc_ctx* ctx);
kc_status kc_graph_download_trace(kc_ctx* ctx, const kc_graph* g, uint64_t* rec, uint32_t cap,
                                  uint32_t* n_out);
uint64_t kc_wall_clock_khz(const kc_ctx* ctx);
/* Per-workgroup loop positions (8 words each) and the control block of the last peel, for
 * diagnosing a KC_ERR_DEVICE time-out. */
kc_status kc_graph_download_debug(kc_ctx* ctx, const kc_graph* g, uint32_t* words, uint32_t cap,
                                  uint32_t* n_out, uint32_t* ctrl_out, uint32_t ctrl_cap);
/* Workgroups of the peel kernel the runtime reports resident per CU, and the CU count. */
kc_status kc_occupancy(kc_ctx* ctx, uint32_t block, int variant, uint32_t* per_cu, uint32_t* cus);
/* Empirical HBM ceiling: the copy kernel that initialises the working degrees (16-byte loads and
 * stores, grid-stride), run `reps` times over two fresh buffers of `bytes` each; *gbps = bytes read
 * plus bytes written per second (10^9), from device events. The buffers are freed before return. */
kc_status kc_copy_bandwidth(kc_ctx* ctx, uint64_t bytes, uint32_t reps, double* gbps);

#ifdef __cplusplus
}
#endif
#endif /* KCORE_H */
