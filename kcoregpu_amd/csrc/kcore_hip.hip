// kcore_hip.hip — host side of libkcore_hip.so: the C ABI of include/kcore.h around the
// gfx950 peeling kernels (peel_kernel.h for one graph, batch_kernel.h for many).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <new>
#include <vector>

#include "../../include/kcore.h"
#include "batch_kernel.h"
#include "host_internal.h"
#include "peel_async.h"
#include "peel_kernel.h"
#include "stepwise_kernel.h"

using namespace kcdev;

namespace {

template <typename T>
__global__ void fill_kernel(T* p, T v, uint64_t n) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t s = (uint64_t)gridDim.x * blockDim.x;
    for (; i < n; i += s) p[i] = v;
}

__global__ void degrees_from_offsets(const uint32_t* off, uint32_t* deg, uint32_t V, uint32_t padV) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t s = gridDim.x * blockDim.x;
    for (; i < padV; i += s) deg[i] = i < V ? off[i + 1] - off[i] : kSent;
}

// working degrees = input degrees (including the kSent padding); a kernel boundary makes
// the copy visible to every XCD before the peel starts
__global__ void copy_kernel(const uint4* __restrict__ src, uint4* __restrict__ dst, uint32_t nvec) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t s = gridDim.x * blockDim.x;
    for (; i < nvec; i += s) dst[i] = src[i];
}

// CSR checks on the device (one pass over the arrays): offsets non-decreasing and every
// neighbour id inside its graph. Sets *bad to 1 on any violation.
__global__ void validate_csr(const uint32_t* off, const uint32_t* nbr, uint32_t V, uint64_t arcs, uint32_t* bad) {
    const uint64_t s = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < arcs; i += s)
        if (nbr[i] >= V) *bad = 1;
    for (uint64_t v = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; v < V; v += s)
        if (off[v + 1] < off[v]) *bad = 1;
}

// one workgroup per graph (grid-strided): offsets start at 0, end at the graph's arc count,
// never decrease, and every local id is below V_g
__global__ void validate_batch(const uint64_t* vtx_off, const uint64_t* arc_off, const uint32_t* off,
                               const uint32_t* nbr, uint32_t G, uint32_t* bad) {
    for (uint32_t g = blockIdx.x; g < G; g += gridDim.x) {
        const uint64_t vb = vtx_off[g], ab = arc_off[g];
        const uint64_t V = vtx_off[g + 1] - vb, A = arc_off[g + 1] - ab;
        const uint32_t* o = off + vb + g;
        if (V >= (1ull << 32) || A >= (1ull << 32) || vtx_off[g + 1] < vb || arc_off[g + 1] < ab) {
            *bad = 1;
            continue;
        }
        if (threadIdx.x == 0 && (o[0] != 0 || o[V] != A)) *bad = 1;
        for (uint64_t v = threadIdx.x; v < V; v += blockDim.x)
            if (o[v + 1] < o[v]) *bad = 1;
        for (uint64_t j = threadIdx.x; j < A; j += blockDim.x)
            if (nbr[ab + j] >= V) *bad = 1;
    }
}

__global__ void batch_degrees(const uint64_t* vtx_off, const uint32_t* off, uint32_t G, uint32_t* deg) {
    for (uint32_t g = blockIdx.x; g < G; g += gridDim.x) {
        const uint64_t vb = vtx_off[g];
        const uint64_t V = vtx_off[g + 1] - vb;
        const uint32_t* o = off + vb + g;
        for (uint64_t v = threadIdx.x; v < V; v += blockDim.x) deg[vb + v] = o[v + 1] - o[v];
    }
}

__global__ void selftest_scan_kernel(const uint32_t* in, uint32_t* out) {
    out[threadIdx.x] = wave_incl_scan(in[threadIdx.x]);
}

constexpr uint32_t kTraceCap = 1u << 16;
constexpr uint32_t kHeavyMin = 16;   // smallest degree that may be processed as chunks
constexpr uint32_t kChunkMin = 64;   // smallest chunk
constexpr uint32_t kMaxGrid = 4096;

}  // namespace

struct kc_graph {
    kc_ctx* ctx = nullptr;
    uint32_t V = 0, padV = 0, dmax = 0;
    uint64_t arcs = 0;
    uint32_t* d_off = nullptr;
    uint32_t* d_nbr = nullptr;
    uint32_t* d_deg0 = nullptr;
    uint32_t* d_core = nullptr;
    uint32_t* d_order = nullptr;
    uint32_t* d_work = nullptr;
    uint64_t* d_chunks = nullptr;
    uint32_t chunk_cap = 0;
    uint32_t* d_shell = nullptr;
    uint32_t* d_wgmin = nullptr;
    uint64_t* d_trace = nullptr;
    uint32_t* d_dbg = nullptr;
    uint32_t* d_ctrl = nullptr;
    uint32_t* d_alive = nullptr;  // two alive lists of padV entries (ASYNC), allocated on first use
    uint32_t chunk_hwm = 0;       // chunk slots written by earlier peels (re-zeroed before ASYNC)
    uint32_t last_grid = 0;
    uint32_t last_kmax = 0;
    uint32_t last_ntrace = 0;
    bool peeled = false;
    bool external = false;        // arrays placed in caller memory (graph_upload_into): never freed
};

// chunks: only a vertex of degree d > heavy (>= kHeavyMin) is cut, into ceil(d / chunk)
// <= d / chunk + 1 pieces, at most once per peel; there are fewer than arcs / kHeavyMin such vertices
static uint32_t chunk_cap_of(uint64_t arcs) {
    return (uint32_t)std::min<uint64_t>(arcs / kHeavyMin + arcs / kChunkMin + 64, 0xFFFFFFF0ull);
}

// The device arrays of a graph (V, padV, arcs, dmax and chunk_cap set), in placement order, with
// their sizes in bytes; the alive lists only when `alive` (caller memory: always placed).
static int graph_arrays(kc_graph* g, bool alive, void** ptr[13], size_t bytes[13]) {
    int n = 0;
    auto add = [&](void* p, size_t b) {
        ptr[n] = static_cast<void**>(p);
        bytes[n] = b;
        ++n;
    };
    add(&g->d_off, ((size_t)g->V + 1) * 4);
    add(&g->d_nbr, std::max<size_t>(g->arcs, 1) * 4);
    add(&g->d_deg0, (size_t)g->padV * 4);
    add(&g->d_core, (size_t)g->padV * 4);
    add(&g->d_order, (size_t)g->padV * 4);
    add(&g->d_work, (size_t)g->padV * 4);
    add(&g->d_chunks, (size_t)g->chunk_cap * 8);
    add(&g->d_shell, ((size_t)g->dmax + 3) * 4);
    add(&g->d_wgmin, (size_t)kMaxGrid * 4);
    add(&g->d_trace, (size_t)kTraceCap * 16);
    add(&g->d_dbg, (size_t)kMaxGrid * 8 * 4);
    add(&g->d_ctrl, (size_t)A_COUNT * kStride * 4);
    if (alive) add(&g->d_alive, (size_t)g->padV * 8);
    return n;
}

static uint64_t align256(uint64_t x) { return (x + 255) & ~255ull; }

// KIND: 0 the sub-round kernel, 1 the cascade-queue kernel, 2 the cascade-queue kernel with the
// level hand-off (no grid barriers)
template <int BLOCK, bool PREFETCH, int KIND>
static kc_status launch_peel(kc_ctx* c, const PeelArgs& args, uint32_t* grid_out) {
    int per_cu = 0;
    if (KIND == 2) {
        KC_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, peel_async_nb<BLOCK, PREFETCH>, BLOCK, 0));
    } else if (KIND == 1) {
        KC_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, peel_async<BLOCK, PREFETCH>, BLOCK, 0));
    } else {
        KC_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, peel_persistent<BLOCK, PREFETCH>, BLOCK, 0));
    }
    if (per_cu < 1) return KC_ERR_HIP;
    uint32_t grid = std::min<uint32_t>((uint32_t)per_cu * (uint32_t)c->cus, kMaxGrid);
    *grid_out = grid;
    PeelArgs a = args;
    if (a.lq_depth > (uint32_t)WgState<BLOCK>::LQ) a.lq_depth = WgState<BLOCK>::LQ;
    // A plain launch: the kernel's elastic start (register_workgroup) only lets workgroups
    // that are resident together take part, so it needs no co-residency guarantee.
    if (KIND == 2) {
        peel_async_nb<BLOCK, PREFETCH><<<dim3(grid), dim3(BLOCK), 0, c->stream>>>(a);
    } else if (KIND == 1) {
        peel_async<BLOCK, PREFETCH><<<dim3(grid), dim3(BLOCK), 0, c->stream>>>(a);
    } else {
        peel_persistent<BLOCK, PREFETCH><<<dim3(grid), dim3(BLOCK), 0, c->stream>>>(a);
    }
    KC_TRY(hipGetLastError());
    return KC_OK;
}

template <int BLOCK>
static kc_status launch_batch(kc_ctx* c, const BatchArgs& a) {
    int per_cu = 0;
    KC_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, peel_batch<BLOCK>, BLOCK, 0));
    if (per_cu < 1) per_cu = 1;
    // graphs are independent: any grid size is correct; more workgroups than graphs is waste
    uint32_t grid = std::min<uint32_t>((uint32_t)per_cu * (uint32_t)c->cus, a.G);
    peel_batch<BLOCK><<<dim3(grid), dim3(BLOCK), 0, c->stream>>>(a);
    KC_TRY(hipGetLastError());
    return KC_OK;
}

// The reference's host-driven level loop over step_select / step_process (see above). Returns
// with core[] holding the core numbers; order[] / shell[] as for the other variants.
static kc_status peel_stepwise(kc_ctx* c, kc_graph* g, uint32_t flags, uint32_t* kmax_out, uint32_t* levels_out,
                               uint32_t* launches_out, float* ms_out) {
    hipStream_t st = c->stream;
    uint32_t* ctr = g->d_ctrl;  // [0] order tail, [1] smallest degree above the level
    struct Pinned {
        uint32_t* p = nullptr;
        ~Pinned() { lib_free_pinned(p); }
    } pin;
    if (const kc_status ps = kc_host_alloc(64, reinterpret_cast<void**>(&pin.p))) return ps;
    uint32_t* h = pin.p;
    std::vector<uint32_t> shell(g->dmax + 3, 0);
    const uint32_t grid = std::min<uint32_t>(4096u, std::max<uint32_t>(1u, (g->V + 255) / 256));
    const uint32_t pgrid = (uint32_t)c->cus * 8;
    kc_status rs = KC_OK;
    uint32_t L = 0, levels = 0, launches = 0, tail = 0;
    KC_TRY(hipEventRecord(c->ev0, st));
    fill_kernel<uint32_t><<<1, 64, 0, st>>>(ctr, 0u, 64);
    copy_kernel<<<2048, 256, 0, st>>>(reinterpret_cast<const uint4*>(g->d_deg0), reinterpret_cast<uint4*>(g->d_core),
                                      g->padV / 4);
    while (true) {
        shell[L] = tail;
        fill_kernel<uint32_t><<<1, 64, 0, st>>>(ctr + 1, kcdev::kSent, 1);
        step_select<<<grid, 256, 0, st>>>(g->d_core, g->V, L, g->d_order, ctr);
        ++launches;
        ++levels;
        uint32_t beg = tail;
        if (hipMemcpyAsync(h, ctr, 8, hipMemcpyDeviceToHost, st) != hipSuccess || hipStreamSynchronize(st) != hipSuccess) {
            rs = KC_ERR_HIP;
            break;
        }
        uint32_t end = h[0];
        while (beg < end) {
            step_process<<<pgrid, 256, 0, st>>>(g->d_off, g->d_nbr, g->d_core, g->d_order, beg, end, L, ctr);
            ++launches;
            if (hipMemcpyAsync(h, ctr, 8, hipMemcpyDeviceToHost, st) != hipSuccess ||
                hipStreamSynchronize(st) != hipSuccess) {
                rs = KC_ERR_HIP;
                break;
            }
            beg = end;
            end = h[0];
        }
        if (rs != KC_OK) break;
        tail = end;
        if (tail >= g->V) {
            shell[L + 1] = tail;
            break;
        }
        uint32_t next = L + 1;
        if (!(flags & KC_FLAG_NO_LEVEL_SKIP) && h[1] != kcdev::kSent && h[1] > next) next = h[1];
        if (next > g->dmax) {
            rs = KC_ERR_NO_PROGRESS;
            break;
        }
        for (uint32_t k = L + 1; k <= next; ++k) shell[k] = tail;
        L = next;
    }
    KC_TRY(hipEventRecord(c->ev1, st));
    KC_TRY(hipEventSynchronize(c->ev1));
    if (rs != KC_OK) return rs;
    KC_TRY(hipGetLastError());
    KC_TRY(hipMemcpyAsync(g->d_shell, shell.data(), (size_t)(L + 2) * 4, hipMemcpyHostToDevice, st));
    KC_TRY(hipStreamSynchronize(st));
    KC_TRY(hipEventElapsedTime(ms_out, c->ev0, c->ev1));
    *kmax_out = L;
    *levels_out = levels;
    *launches_out = launches;
    return KC_OK;
}

// Builds a resident graph: its arrays are allocated one by one (mem == nullptr) or placed in
// [mem, mem + bytes) at 256-byte boundaries (external); the CSR is copied in and checked on the
// device before any peeling kernel may index with it.
static kc_status upload_impl(kc_ctx* c, const kc_csr* h, uint8_t* mem, uint64_t bytes, kc_graph** out) {
    if (!c || !h || !out || !h->offsets || (h->arcs && !h->neighbors)) return KC_ERR_ARG;
    if (mem && (reinterpret_cast<uintptr_t>(mem) & 255u)) return KC_ERR_ARG;
    // padV = V rounded up to 256 must keep the degree array below 4 GiB (one buffer descriptor)
    if (h->V == 0 || h->V > (1u << 30) - 256 || h->arcs >= 0xFFFFFFFFull) return KC_ERR_CAPACITY;
    if (h->offsets[0] != 0 || h->offsets[h->V] != h->arcs) return KC_ERR_ARG;
    KC_TRY(hipSetDevice(c->device));
    kc_graph* g = new (std::nothrow) kc_graph;
    if (!g) return KC_ERR_OOM;
    g->ctx = c;
    g->V = h->V;
    g->arcs = h->arcs;
    g->padV = (h->V + 255u) & ~255u;
    uint32_t dmax = 0;
    for (uint32_t v = 0; v < h->V; ++v) {
        if (h->offsets[v + 1] < h->offsets[v]) {
            delete g;
            return KC_ERR_ARG;
        }
        dmax = std::max(dmax, h->offsets[v + 1] - h->offsets[v]);
    }
    g->dmax = dmax;
    g->chunk_cap = chunk_cap_of(h->arcs);
    g->external = mem != nullptr;
    auto fail = [&](hipError_t e) {
        kc_graph_free(g);
        return hip_status(e);
    };
    void** ptr[13];
    size_t sz[13];
    const int n = graph_arrays(g, g->external, ptr, sz);
    hipError_t e;
    if (g->external) {
        uint64_t need = 0;
        for (int i = 0; i < n; ++i) need += align256(sz[i]);
        if (need > bytes) {
            delete g;
            return KC_ERR_CAPACITY;
        }
        uint64_t pos = 0;
        for (int i = 0; i < n; ++i) {
            *ptr[i] = mem + pos;
            pos += align256(sz[i]);
        }
    } else {
        for (int i = 0; i < n; ++i)
            if ((e = hipMalloc(ptr[i], sz[i])) != hipSuccess) return fail(e);
    }
    if ((e = hipMemcpyAsync(g->d_off, h->offsets, ((size_t)h->V + 1) * 4, hipMemcpyHostToDevice, c->stream)) != hipSuccess) return fail(e);
    if (h->arcs && (e = hipMemcpyAsync(g->d_nbr, h->neighbors, h->arcs * 4, hipMemcpyHostToDevice, c->stream)) != hipSuccess) return fail(e);
    if (h->degrees) {
        if ((e = hipMemcpyAsync(g->d_deg0, h->degrees, (size_t)h->V * 4, hipMemcpyHostToDevice, c->stream)) != hipSuccess) return fail(e);
        if (g->padV > h->V) fill_kernel<uint32_t><<<1, 256, 0, c->stream>>>(g->d_deg0 + h->V, kSent, g->padV - h->V);
    } else {
        degrees_from_offsets<<<1024, 256, 0, c->stream>>>(g->d_off, g->d_deg0, g->V, g->padV);
    }
    // chunk slots start empty (0): the ASYNC kernel waits on a claimed slot until it is written
    fill_kernel<uint64_t><<<1024, 256, 0, c->stream>>>(g->d_chunks, 0ull, g->chunk_cap);
    // reject a malformed CSR before any kernel indexes with it
    fill_kernel<uint32_t><<<1, 64, 0, c->stream>>>(g->d_ctrl, 0u, 1);
    validate_csr<<<1024, 256, 0, c->stream>>>(g->d_off, g->d_nbr, g->V, g->arcs, g->d_ctrl);
    if ((e = hipGetLastError()) != hipSuccess) return fail(e);
    uint32_t bad = 0;
    if ((e = hipMemcpyAsync(&bad, g->d_ctrl, 4, hipMemcpyDeviceToHost, c->stream)) != hipSuccess) return fail(e);
    if ((e = hipStreamSynchronize(c->stream)) != hipSuccess) return fail(e);
    if (bad) {
        kc_graph_free(g);
        return KC_ERR_ARG;
    }
    *out = g;
    return KC_OK;
}

uint64_t graph_bytes(uint32_t V, uint64_t arcs, uint32_t dmax) {
    kc_graph g;
    g.V = V;
    g.padV = (V + 255u) & ~255u;
    g.arcs = arcs;
    g.dmax = dmax;
    g.chunk_cap = chunk_cap_of(arcs);
    void** ptr[13];
    size_t sz[13];
    const int n = graph_arrays(&g, true, ptr, sz);
    uint64_t need = 0;
    for (int i = 0; i < n; ++i) need += align256(sz[i]);
    return need;
}

kc_status graph_upload_into(kc_ctx* c, const kc_csr* h, void* mem, uint64_t bytes, kc_graph** out) {
    if (!mem) return KC_ERR_ARG;
    return upload_impl(c, h, static_cast<uint8_t*>(mem), bytes, out);
}

extern "C" {

const char* kc_status_str(kc_status s) {
    switch (s) {
        case KC_OK: return "ok";
        case KC_ERR_ARG: return "invalid argument";
        case KC_ERR_HIP: return "HIP runtime error";
        case KC_ERR_OOM: return "device out of memory";
        case KC_ERR_CAPACITY: return "request exceeds a documented capacity";
        case KC_ERR_NO_PROGRESS: return "level loop made no progress";
        case KC_ERR_DEVICE: return "device-side wait exceeded its time limit";
        case KC_ERR_NO_DEVICE: return "no usable GPU";
        default: return "unknown status";
    }
}

kc_status kc_open(int device, kc_ctx** out) {
    if (!out) return KC_ERR_ARG;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= device || device < 0) return KC_ERR_NO_DEVICE;
    KC_TRY(hipSetDevice(device));
    kc_ctx* c = new (std::nothrow) kc_ctx;
    if (!c) return KC_ERR_OOM;
    c->device = device;
    hipDeviceProp_t prop;
    KC_TRY(hipGetDeviceProperties(&prop, device));
    c->cus = prop.multiProcessorCount;
    std::snprintf(c->arch, sizeof(c->arch), "%s", prop.gcnArchName);
    for (char* q = c->arch; *q; ++q)
        if (*q == ' ') *q = '_';
    int rate = 0;
    if (hipDeviceGetAttribute(&rate, hipDeviceAttributeWallClockRate, device) == hipSuccess && rate > 0)
        c->wall_rate_khz = (uint64_t)rate;
    if (const char* e = std::getenv("KCORE_WAIT_LIMIT_MS")) {
        const long long v = std::atoll(e);
        if (v > 0) c->wait_limit_ms = (uint64_t)v;
    }
    KC_TRY(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    KC_TRY(hipEventCreate(&c->ev0));
    KC_TRY(hipEventCreate(&c->ev1));
    KC_TRY(hipEventCreate(&c->evk0));
    KC_TRY(hipEventCreate(&c->evk1));
    *out = c;
    return KC_OK;
}

void kc_close(kc_ctx* c) {
    if (!c) return;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    for (hipEvent_t e : {c->ev0, c->ev1, c->evk0, c->evk1})
        if (e) (void)hipEventDestroy(e);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

void kc_graph_free(kc_graph* g) {
    if (!g) return;
    (void)hipSetDevice(g->ctx->device);
    (void)hipStreamSynchronize(g->ctx->stream);
    if (!g->external) {
        void** ptr[13];
        size_t sz[13];
        const int n = graph_arrays(g, true, ptr, sz);
        for (int i = 0; i < n; ++i) lib_free_device(*ptr[i]);
    }
    delete g;
}

kc_status kc_graph_upload(kc_ctx* c, const kc_csr* h, kc_graph** out) {
    return upload_impl(c, h, nullptr, 0, out);
}

void kc_graph_info(const kc_graph* g, uint32_t* V, uint64_t* arcs, uint32_t* dmax) {
    if (V) *V = g->V;
    if (arcs) *arcs = g->arcs;
    if (dmax) *dmax = g->dmax;
}

kc_status kc_peel(kc_ctx* c, kc_graph* g, const kc_params* p, uint32_t* core_out,
                  uint32_t* kmax_out, kc_stats* st) {
    if (!c || !g || g->ctx != c) return KC_ERR_ARG;
    KC_TRY(hipSetDevice(c->device));
    kc_params prm{};
    if (p) prm = *p;
    int variant = prm.variant ? prm.variant : KC_VARIANT_LDS;
    if (variant < KC_VARIANT_WAVE || variant > KC_VARIANT_ASYNC) return KC_ERR_ARG;
    if (variant == KC_VARIANT_STEPWISE) {
        uint32_t kmax = 0, levels = 0, launches = 0;
        float ms = 0;
        const kc_status s = peel_stepwise(c, g, prm.flags, &kmax, &levels, &launches, &ms);
        if (s != KC_OK) return s;
        g->last_kmax = kmax;
        g->last_ntrace = 0;
        g->peeled = true;
        if (core_out) {
            KC_TRY(hipMemcpyAsync(core_out, g->d_core, (size_t)g->V * 4, hipMemcpyDeviceToHost, c->stream));
            KC_TRY(hipStreamSynchronize(c->stream));
        }
        if (kmax_out) *kmax_out = kmax;
        if (st) {
            memset(st, 0, sizeof(*st));
            st->ms_peel = ms;
            st->ms_kernel = ms;
            st->kmax = kmax;
            st->levels = levels;
            st->subrounds = launches;
            st->scanned_entries = (uint64_t)levels * g->V;
            st->bytes_scan = 4 * st->scanned_entries;
            st->bytes_edge = 16ull * g->V + 20ull * (g->arcs / 2);
            st->bytes_alg = st->bytes_scan + st->bytes_edge;
            st->block = 256;
        }
        return KC_OK;
    }
    const bool async = variant == KC_VARIANT_ASYNC;
    if (async && !g->d_alive) KC_TRY(hipMalloc(&g->d_alive, (size_t)g->padV * 8));
    const uint32_t block = prm.block ? prm.block : 512;
    if (block != 256 && block != 512 && block != 1024) return KC_ERR_ARG;
    uint32_t heavy = prm.heavy_threshold ? prm.heavy_threshold : 64;
    if (heavy < kHeavyMin) heavy = kHeavyMin;
    uint32_t chunk = prm.chunk_arcs ? prm.chunk_arcs : 512;
    if (chunk < kChunkMin) chunk = kChunkMin;

    PeelArgs a{};
    a.off = g->d_off;
    a.nbr = g->d_nbr;
    a.deg0 = g->d_deg0;
    a.core = g->d_core;
    a.order = g->d_order;
    a.work = g->d_work;
    a.chunks = g->d_chunks;
    a.shell = g->d_shell;
    a.wgmin = g->d_wgmin;
    a.trace = g->d_trace;
    a.dbg = g->d_dbg;
    a.ctrl = g->d_ctrl;
    a.V = g->V;
    a.nvec = g->padV / 4;
    a.dmax = g->dmax;
    a.heavy = heavy;
    a.chunk = chunk;
    a.chunk_cap = g->chunk_cap;
    a.lq_depth = variant == KC_VARIANT_WAVE ? 0 : (prm.lds_depth ? prm.lds_depth : 0xFFFFFFFFu);
    a.lq_arcs = 16384;
    a.lq_rounds = 4;
    a.flags = prm.flags;
    a.trace_cap = kTraceCap;
    a.time_limit = c->wall_rate_khz * c->wait_limit_ms;
    a.reg_window = c->wall_rate_khz * 30 / 1000;  // 30 us without a new check-in
    a.alive[0] = g->d_alive;
    a.alive[1] = g->d_alive ? g->d_alive + g->padV : nullptr;
    a.async_mode = async ? 1u : 0u;
    a.compact_ratio = prm.compact_ratio == 0xFFFFFFFFu ? 0u : (prm.compact_ratio ? prm.compact_ratio : 4u);
    a.poll_max = prm.poll_max ? std::min<uint32_t>(prm.poll_max, 3u) : 2u;

    KC_TRY(hipEventRecord(c->ev0, c->stream));
    fill_kernel<uint32_t><<<256, 256, 0, c->stream>>>(g->d_ctrl, 0u, (uint64_t)A_COUNT * kStride);
    if (async) {
        // queue slots start as "not yet written": order[] ~0, chunk slots 0
        fill_kernel<uint32_t><<<1024, 256, 0, c->stream>>>(g->d_order, kSent, g->V);
        if (g->chunk_hwm)
            fill_kernel<uint64_t><<<1024, 256, 0, c->stream>>>(g->d_chunks, 0ull, g->chunk_hwm);
    }
    fill_kernel<uint32_t><<<64, 256, 0, c->stream>>>(g->d_shell, 0u, (uint64_t)g->dmax + 3);
    fill_kernel<uint32_t><<<64, 256, 0, c->stream>>>(g->d_dbg, 0u, (uint64_t)kMaxGrid * 8);
    KC_TRY(hipGetLastError());
    uint32_t grid = 0;
    kc_status s;
    KC_TRY(hipEventRecord(c->evk0, c->stream));
    copy_kernel<<<2048, 256, 0, c->stream>>>(reinterpret_cast<const uint4*>(g->d_deg0),
                                             reinterpret_cast<uint4*>(g->d_core), g->padV / 4);
    KC_TRY(hipGetLastError());
    const bool pf = variant == KC_VARIANT_PREFETCH;
    if (async && (prm.flags & KC_FLAG_LEVEL_HANDOFF)) {
        switch (block) {
            case 256: s = launch_peel<256, false, 2>(c, a, &grid); break;
            case 512: s = launch_peel<512, false, 2>(c, a, &grid); break;
            default: s = launch_peel<1024, false, 2>(c, a, &grid); break;
        }
    } else if (async) {
        switch (block) {
            case 256: s = launch_peel<256, false, 1>(c, a, &grid); break;
            case 512: s = launch_peel<512, false, 1>(c, a, &grid); break;
            default: s = launch_peel<1024, false, 1>(c, a, &grid); break;
        }
    } else {
        switch (block) {
            case 256: s = pf ? launch_peel<256, true, 0>(c, a, &grid) : launch_peel<256, false, 0>(c, a, &grid); break;
            case 512: s = pf ? launch_peel<512, true, 0>(c, a, &grid) : launch_peel<512, false, 0>(c, a, &grid); break;
            default: s = pf ? launch_peel<1024, true, 0>(c, a, &grid) : launch_peel<1024, false, 0>(c, a, &grid); break;
        }
    }
    if (s != KC_OK) return s;
    g->last_grid = grid;
    KC_TRY(hipEventRecord(c->evk1, c->stream));
    KC_TRY(hipEventRecord(c->ev1, c->stream));
    uint32_t res[R_WORDS + 4];
    KC_TRY(hipMemcpyAsync(res, g->d_ctrl + (size_t)C_RESULT * kStride, sizeof(res), hipMemcpyDeviceToHost, c->stream));
    KC_TRY(hipStreamSynchronize(c->stream));
    float ms = 0, msk = 0;
    KC_TRY(hipEventElapsedTime(&ms, c->ev0, c->ev1));
    KC_TRY(hipEventElapsedTime(&msk, c->evk0, c->evk1));
    const uint32_t err = res[R_ERR];
    if (err == 1) return KC_ERR_DEVICE;
    if (err == 2) return KC_ERR_NO_PROGRESS;
    if (res[R_NPART] == 0) return KC_ERR_HIP;  // the kernel never ran
    g->last_kmax = res[R_KMAX];
    g->chunk_hwm = std::max(g->chunk_hwm, std::min(res[R_WORDS + 3], g->chunk_cap));
    g->last_ntrace = std::min(res[R_NTRACE], kTraceCap);
    g->peeled = true;
    double d2h = 0;
    if (core_out) {
        KC_TRY(hipEventRecord(c->ev0, c->stream));
        KC_TRY(hipMemcpyAsync(core_out, g->d_core, (size_t)g->V * 4, hipMemcpyDeviceToHost, c->stream));
        KC_TRY(hipEventRecord(c->ev1, c->stream));
        KC_TRY(hipStreamSynchronize(c->stream));
        float m2 = 0;
        KC_TRY(hipEventElapsedTime(&m2, c->ev0, c->ev1));
        d2h = m2;
    }
    if (kmax_out) *kmax_out = g->last_kmax;
    if (st) {
        memset(st, 0, sizeof(*st));
        st->ms_peel = ms;
        st->ms_kernel = msk;
        st->ms_d2h = d2h;
        st->kmax = g->last_kmax;
        st->levels = res[R_LEVELS];
        st->subrounds = res[R_SUBROUNDS];
        if (async) {
            // dense entries (4 B), list entries (4 B id + 4 B gathered degree), list writes (4 B)
            st->scanned_entries = (uint64_t)res[R_WORDS] + res[R_WORDS + 1];
            st->bytes_scan = 4ull * res[R_WORDS] + 8ull * res[R_WORDS + 1] + 4ull * res[R_WORDS + 2];
        } else {
            st->scanned_entries = (uint64_t)st->levels * g->padV;
            st->bytes_scan = 4 * st->scanned_entries;
        }
        st->bytes_edge = 16ull * g->V + 20ull * (g->arcs / 2);
        st->bytes_alg = st->bytes_scan + st->bytes_edge;
        st->grid = grid;
        st->block = block;
        st->participants = res[R_NPART];
    }
    return KC_OK;
}

kc_status kc_graph_download_core(kc_ctx* c, const kc_graph* g, uint32_t* core_out) {
    if (!c || !g || !core_out || !g->peeled) return KC_ERR_ARG;
    KC_TRY(hipSetDevice(c->device));
    KC_TRY(hipMemcpyAsync(core_out, g->d_core, (size_t)g->V * 4, hipMemcpyDeviceToHost, c->stream));
    KC_TRY(hipStreamSynchronize(c->stream));
    return KC_OK;
}

kc_status kc_graph_download_order(kc_ctx* c, const kc_graph* g, uint32_t* order_out,
                                  uint32_t* shell_off_out) {
    if (!c || !g || !g->peeled) return KC_ERR_ARG;
    KC_TRY(hipSetDevice(c->device));
    if (order_out)
        KC_TRY(hipMemcpyAsync(order_out, g->d_order, (size_t)g->V * 4, hipMemcpyDeviceToHost, c->stream));
    if (shell_off_out)
        KC_TRY(hipMemcpyAsync(shell_off_out, g->d_shell, ((size_t)g->last_kmax + 2) * 4,
                              hipMemcpyDeviceToHost, c->stream));
    KC_TRY(hipStreamSynchronize(c->stream));
    return KC_OK;
}

kc_status kc_graph_download_trace(kc_ctx* c, const kc_graph* g, uint64_t* rec, uint32_t cap,
                                  uint32_t* n_out) {
    if (!c || !g || !g->peeled || !n_out) return KC_ERR_ARG;
    KC_TRY(hipSetDevice(c->device));
    const uint32_t n = std::min(cap, g->last_ntrace);
    if (n && rec)
        KC_TRY(hipMemcpyAsync(rec, g->d_trace, (size_t)n * 16, hipMemcpyDeviceToHost, c->stream));
    KC_TRY(hipStreamSynchronize(c->stream));
    *n_out = n;
    return KC_OK;
}

const uint32_t* kc_graph_core_device(const kc_graph* g) { return g ? g->d_core : nullptr; }

kc_status kc_graph_download_debug(kc_ctx* c, const kc_graph* g, uint32_t* words, uint32_t cap,
                                  uint32_t* n_out, uint32_t* ctrl_out, uint32_t ctrl_cap) {
    if (!c || !g || !n_out) return KC_ERR_ARG;
    KC_TRY(hipSetDevice(c->device));
    const uint32_t n = std::min(cap, g->last_grid * 8);
    if (n && words) KC_TRY(hipMemcpyAsync(words, g->d_dbg, (size_t)n * 4, hipMemcpyDeviceToHost, c->stream));
    const uint32_t nc = std::min<uint32_t>(ctrl_cap, C_COUNT * kStride);
    if (nc && ctrl_out) KC_TRY(hipMemcpyAsync(ctrl_out, g->d_ctrl, (size_t)nc * 4, hipMemcpyDeviceToHost, c->stream));
    KC_TRY(hipStreamSynchronize(c->stream));
    *n_out = n;
    return KC_OK;
}

uint64_t kc_wall_clock_khz(const kc_ctx* c) { return c ? c->wall_rate_khz : 0; }

kc_status kc_occupancy(kc_ctx* c, uint32_t block, int variant, uint32_t* per_cu, uint32_t* cus) {
    if (!c || !per_cu || !cus) return KC_ERR_ARG;
    KC_TRY(hipSetDevice(c->device));
    int n = 0;
    const bool pf = variant == KC_VARIANT_PREFETCH || variant == KC_VARIANT_AUTO;
    switch (block) {
        case 256: KC_TRY(pf ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, peel_persistent<256, true>, 256, 0)
                            : hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, peel_persistent<256, false>, 256, 0)); break;
        case 512: KC_TRY(pf ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, peel_persistent<512, true>, 512, 0)
                            : hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, peel_persistent<512, false>, 512, 0)); break;
        case 1024: KC_TRY(pf ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, peel_persistent<1024, true>, 1024, 0)
                             : hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, peel_persistent<1024, false>, 1024, 0)); break;
        default: return KC_ERR_ARG;
    }
    *per_cu = (uint32_t)n;
    *cus = (uint32_t)c->cus;
    return KC_OK;
}

kc_status kc_copy_bandwidth(kc_ctx* c, uint64_t bytes, uint32_t reps, double* gbps) {
    if (!c || !gbps || bytes < 16 || reps == 0) return KC_ERR_ARG;
    const uint64_t nvec = bytes / 16;
    if (nvec > 0xFFFFFFFFull) return KC_ERR_CAPACITY;
    KC_TRY(hipSetDevice(c->device));
    uint4 *src = nullptr, *dst = nullptr;
    KC_TRY(hipMalloc(&src, nvec * 16));
    if (hipMalloc(&dst, nvec * 16) != hipSuccess) {
        lib_free_device(src);
        return KC_ERR_OOM;
    }
    const uint32_t grid = (uint32_t)c->cus * 16;
    fill_kernel<uint32_t><<<1024, 256, 0, c->stream>>>(reinterpret_cast<uint32_t*>(src), 7u, nvec * 4);
    copy_kernel<<<grid, 256, 0, c->stream>>>(src, dst, (uint32_t)nvec);  // warm-up
    hipError_t e = hipEventRecord(c->ev0, c->stream);
    for (uint32_t r = 0; r < reps && e == hipSuccess; ++r) copy_kernel<<<grid, 256, 0, c->stream>>>(src, dst, (uint32_t)nvec);
    if (e == hipSuccess) e = hipGetLastError();
    if (e == hipSuccess) e = hipEventRecord(c->ev1, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    float ms = 0;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, c->ev0, c->ev1);
    lib_free_device(src);
    lib_free_device(dst);
    KC_TRY(e);
    *gbps = ms > 0 ? 2.0 * (double)nvec * 16.0 * reps / (ms * 1e-3) / 1e9 : 0.0;
    return KC_OK;
}

kc_status kc_selftest(kc_ctx* c) {
    if (!c) return KC_ERR_ARG;
    KC_TRY(hipSetDevice(c->device));
    uint32_t h_in[64], h_out[64];
    for (int i = 0; i < 64; ++i) h_in[i] = (uint32_t)(i * 7 + 3) % 11 + (i == 40 ? 100000u : 0u);
    uint32_t *d_in = nullptr, *d_out = nullptr;
    KC_TRY(hipMalloc(&d_in, 256));
    if (hipMalloc(&d_out, 256) != hipSuccess) {
        lib_free_device(d_in);
        return KC_ERR_OOM;
    }
    hipError_t e = hipMemcpyAsync(d_in, h_in, 256, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) {
        selftest_scan_kernel<<<1, 64, 0, c->stream>>>(d_in, d_out);
        e = hipMemcpyAsync(h_out, d_out, 256, hipMemcpyDeviceToHost, c->stream);
    }
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    lib_free_device(d_in);
    lib_free_device(d_out);
    KC_TRY(e);
    uint32_t acc = 0;
    for (int i = 0; i < 64; ++i) {
        acc += h_in[i];
        if (h_out[i] != acc) return KC_ERR_DEVICE;
    }
    return KC_OK;
}


// ---------------------------------------------------------------------------------------
// batches

struct kc_batch {
    kc_ctx* ctx = nullptr;
    uint32_t G = 0;
    uint64_t nv = 0, na = 0, padv = 0;
    uint64_t *d_vtx = nullptr, *d_arc = nullptr;
    uint32_t *d_off = nullptr, *d_nbr = nullptr, *d_deg0 = nullptr, *d_core = nullptr, *d_order = nullptr;
    uint32_t *d_kmax = nullptr, *d_sched = nullptr, *d_ctrl = nullptr;
    std::vector<uint64_t> vtx_off, arc_off;
    bool peeled = false;
    struct Captured {
        uint32_t block, flags;
        hipGraph_t graph;
        hipGraphExec_t exec;
    };
    std::vector<Captured> graphs;  // captured launch sequences (KC_FLAG_GRAPH)
};

void kc_batch_free(kc_batch* b) {
    if (!b) return;
    (void)hipSetDevice(b->ctx->device);
    (void)hipStreamSynchronize(b->ctx->stream);
    for (void* p : {(void*)b->d_vtx, (void*)b->d_arc, (void*)b->d_off, (void*)b->d_nbr, (void*)b->d_deg0,
                    (void*)b->d_core, (void*)b->d_order, (void*)b->d_kmax, (void*)b->d_sched, (void*)b->d_ctrl})
        lib_free_device(p);
    for (auto& cg : b->graphs) {
        (void)hipGraphExecDestroy(cg.exec);
        (void)hipGraphDestroy(cg.graph);
    }
    delete b;
}

kc_status kc_batch_upload(kc_ctx* c, const kc_batch_csr* h, kc_batch** out) {
    if (!c || !h || !out || h->G == 0 || !h->vtx_off || !h->arc_off || !h->offsets ||
        (h->total_arcs && !h->neighbors))
        return KC_ERR_ARG;
    if (h->vtx_off[0] != 0 || h->arc_off[0] != 0 || h->vtx_off[h->G] != h->total_vertices ||
        h->arc_off[h->G] != h->total_arcs)
        return KC_ERR_ARG;
    // graph ranges must be non-decreasing before any device code indexes with them; graph
    // sizes must fit the kernel's 32-bit per-graph indices
    for (uint32_t g = 0; g < h->G; ++g) {
        if (h->vtx_off[g + 1] < h->vtx_off[g] || h->arc_off[g + 1] < h->arc_off[g]) return KC_ERR_ARG;
        if (h->vtx_off[g + 1] - h->vtx_off[g] >= (1ull << 30) || h->arc_off[g + 1] - h->arc_off[g] >= (1ull << 32))
            return KC_ERR_CAPACITY;
    }
    KC_TRY(hipSetDevice(c->device));
    kc_batch* b = new (std::nothrow) kc_batch;
    if (!b) return KC_ERR_OOM;
    b->ctx = c;
    b->G = h->G;
    b->nv = h->total_vertices;
    b->na = h->total_arcs;
    b->padv = (b->nv + 255) & ~255ull;
    b->vtx_off.assign(h->vtx_off, h->vtx_off + h->G + 1);
    b->arc_off.assign(h->arc_off, h->arc_off + h->G + 1);
    // claim order: most arcs first, so the longest graphs start first
    std::vector<uint32_t> sched(h->G);
    for (uint32_t g = 0; g < h->G; ++g) sched[g] = g;
    std::stable_sort(sched.begin(), sched.end(), [&](uint32_t x, uint32_t y) {
        return h->arc_off[x + 1] - h->arc_off[x] > h->arc_off[y + 1] - h->arc_off[y];
    });
    auto fail = [&](hipError_t e) {
        kc_batch_free(b);
        return hip_status(e);
    };
    hipError_t e;
#define KC_ALLOC(ptr, bytes) \
    if ((e = hipMalloc(&(ptr), (bytes))) != hipSuccess) return fail(e)
    KC_ALLOC(b->d_vtx, ((size_t)h->G + 1) * 8);
    KC_ALLOC(b->d_arc, ((size_t)h->G + 1) * 8);
    KC_ALLOC(b->d_off, (b->nv + h->G) * 4);
    KC_ALLOC(b->d_nbr, std::max<uint64_t>(b->na, 1) * 4);
    KC_ALLOC(b->d_deg0, b->padv * 4);
    KC_ALLOC(b->d_core, b->padv * 4);
    KC_ALLOC(b->d_order, b->padv * 4);
    KC_ALLOC(b->d_kmax, (size_t)h->G * 4);
    KC_ALLOC(b->d_sched, (size_t)h->G * 4);
    KC_ALLOC(b->d_ctrl, 4096);
#undef KC_ALLOC
    hipStream_t st = c->stream;
    if ((e = hipMemcpyAsync(b->d_vtx, h->vtx_off, ((size_t)h->G + 1) * 8, hipMemcpyHostToDevice, st)) ||
        (e = hipMemcpyAsync(b->d_arc, h->arc_off, ((size_t)h->G + 1) * 8, hipMemcpyHostToDevice, st)) ||
        (e = hipMemcpyAsync(b->d_off, h->offsets, (b->nv + h->G) * 4, hipMemcpyHostToDevice, st)) ||
        (e = hipMemcpyAsync(b->d_sched, sched.data(), (size_t)h->G * 4, hipMemcpyHostToDevice, st)))
        return fail(e);
    if (b->na && (e = hipMemcpyAsync(b->d_nbr, h->neighbors, b->na * 4, hipMemcpyHostToDevice, st))) return fail(e);
    fill_kernel<uint32_t><<<1, 64, 0, st>>>(b->d_ctrl, 0u, 1024);
    validate_batch<<<1024, 256, 0, st>>>(b->d_vtx, b->d_arc, b->d_off, b->d_nbr, h->G, b->d_ctrl);
    fill_kernel<uint32_t><<<64, 256, 0, st>>>(b->d_deg0, kSent, b->padv);
    if (h->degrees) {
        if ((e = hipMemcpyAsync(b->d_deg0, h->degrees, b->nv * 4, hipMemcpyHostToDevice, st))) return fail(e);
    } else {
        batch_degrees<<<1024, 256, 0, st>>>(b->d_vtx, b->d_off, h->G, b->d_deg0);
    }
    if ((e = hipGetLastError()) != hipSuccess) return fail(e);
    uint32_t bad = 0;
    if ((e = hipMemcpyAsync(&bad, b->d_ctrl, 4, hipMemcpyDeviceToHost, st)) != hipSuccess) return fail(e);
    if ((e = hipStreamSynchronize(st)) != hipSuccess) return fail(e);
    if (bad) {
        kc_batch_free(b);
        return KC_ERR_ARG;
    }
    *out = b;
    return KC_OK;
}

kc_status kc_batch_peel(kc_ctx* c, kc_batch* b, const kc_params* p, uint32_t* core_out, uint32_t* kmax_out,
                        kc_stats* stt) {
    if (!c || !b || b->ctx != c) return KC_ERR_ARG;
    KC_TRY(hipSetDevice(c->device));
    kc_params prm{};
    if (p) prm = *p;
    const uint32_t block = prm.block ? prm.block : 512;
    if (block != 256 && block != 512 && block != 1024) return KC_ERR_ARG;
    BatchArgs a{};
    a.vtx_off = b->d_vtx;
    a.arc_off = b->d_arc;
    a.off = b->d_off;
    a.nbr = b->d_nbr;
    a.core = b->d_core;
    a.order = b->d_order;
    a.kmax = b->d_kmax;
    a.sched = b->d_sched;
    a.ctrl = b->d_ctrl;
    a.G = b->G;
    a.flags = prm.flags;
    // the launch sequence: state reset (a fill kernel, not a memset node), degree copy, peel
    auto enqueue = [&](bool timed) -> kc_status {
        fill_kernel<uint32_t><<<1, 256, 0, c->stream>>>(b->d_ctrl, 0u, 1024);
        copy_kernel<<<2048, 256, 0, c->stream>>>(reinterpret_cast<const uint4*>(b->d_deg0),
                                                 reinterpret_cast<uint4*>(b->d_core), (uint32_t)(b->padv / 4));
        KC_TRY(hipGetLastError());
        if (timed) KC_TRY(hipEventRecord(c->evk0, c->stream));
        kc_status s;
        switch (block) {
            case 256: s = launch_batch<256>(c, a); break;
            case 512: s = launch_batch<512>(c, a); break;
            default: s = launch_batch<1024>(c, a); break;
        }
        if (s != KC_OK) return s;
        if (timed) KC_TRY(hipEventRecord(c->evk1, c->stream));
        return KC_OK;
    };
    const bool use_graph = (prm.flags & KC_FLAG_GRAPH) != 0;
    hipGraphExec_t exec = nullptr;
    if (use_graph) {
        for (auto& cg : b->graphs)
            if (cg.block == block && cg.flags == prm.flags) exec = cg.exec;
        if (!exec) {
            kc_batch::Captured cg{block, prm.flags, nullptr, nullptr};
            KC_TRY(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
            const kc_status s = enqueue(false);
            const hipError_t ce = hipStreamEndCapture(c->stream, &cg.graph);
            if (s != KC_OK) {
                if (cg.graph) (void)hipGraphDestroy(cg.graph);
                return s;
            }
            KC_TRY(ce);
            const hipError_t ie = hipGraphInstantiate(&cg.exec, cg.graph, nullptr, nullptr, 0);
            if (ie != hipSuccess) {
                (void)hipGraphDestroy(cg.graph);
                return hip_status(ie);
            }
            b->graphs.push_back(cg);
            exec = cg.exec;
        }
    }
    KC_TRY(hipEventRecord(c->ev0, c->stream));
    if (use_graph) {
        KC_TRY(hipGraphLaunch(exec, c->stream));
    } else {
        const kc_status s = enqueue(true);
        if (s != KC_OK) return s;
    }
    KC_TRY(hipEventRecord(c->ev1, c->stream));
    uint32_t ctl[256];
    KC_TRY(hipMemcpyAsync(ctl, b->d_ctrl, sizeof(ctl), hipMemcpyDeviceToHost, c->stream));
    KC_TRY(hipStreamSynchronize(c->stream));
    if (ctl[192]) return KC_ERR_NO_PROGRESS;
    float ms = 0, msk = 0;
    KC_TRY(hipEventElapsedTime(&ms, c->ev0, c->ev1));
    if (use_graph) {
        msk = ms;
    } else {
        KC_TRY(hipEventElapsedTime(&msk, c->evk0, c->evk1));
    }
    b->peeled = true;
    if (core_out && b->nv)
        KC_TRY(hipMemcpyAsync(core_out, b->d_core, b->nv * 4, hipMemcpyDeviceToHost, c->stream));
    if (kmax_out) KC_TRY(hipMemcpyAsync(kmax_out, b->d_kmax, (size_t)b->G * 4, hipMemcpyDeviceToHost, c->stream));
    KC_TRY(hipStreamSynchronize(c->stream));
    if (stt) {
        memset(stt, 0, sizeof(*stt));
        stt->ms_peel = ms;
        stt->ms_kernel = msk;
        stt->levels = ctl[64];
        stt->subrounds = ctl[128];
        stt->block = block;
        // scan term: each graph scans its V_g entries once per level it runs
        stt->scanned_entries = (uint64_t)ctl[224] | ((uint64_t)ctl[225] << 32);
        stt->bytes_scan = 4 * stt->scanned_entries;
        stt->bytes_edge = 16ull * b->nv + 20ull * (b->na / 2);
        stt->bytes_alg = stt->bytes_scan + stt->bytes_edge;
    }
    return KC_OK;
}

void kc_batch_info(const kc_batch* b, uint32_t* G, uint64_t* total_vertices, uint64_t* total_arcs) {
    if (G) *G = b->G;
    if (total_vertices) *total_vertices = b->nv;
    if (total_arcs) *total_arcs = b->na;
}

kc_status kc_batch_download_order(kc_ctx* c, const kc_batch* b, uint32_t* order_out) {
    if (!c || !b || !order_out || !b->peeled) return KC_ERR_ARG;
    KC_TRY(hipSetDevice(c->device));
    if (b->nv) KC_TRY(hipMemcpyAsync(order_out, b->d_order, b->nv * 4, hipMemcpyDeviceToHost, c->stream));
    KC_TRY(hipStreamSynchronize(c->stream));
    return KC_OK;
}

}  // extern "C"
