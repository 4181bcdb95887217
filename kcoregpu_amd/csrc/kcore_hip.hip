// kcore_hip.hip — MI355X (gfx950, CDNA4, wave64) k-core peeling engine and its C ABI
// (include/kcore.h).
//
// Method: the reference's level-synchronous peel (src/ours.cc:2-160). For level L:
//   select  every v with deg[v] == L joins the frontier            (src/ours.cc:2-33)
//   process for each frontier v and neighbour u with deg[u] > L:
//           a = atomicSub(&deg[u], 1); a == L+1 -> u joins the frontier of this level;
//           a <= L -> undo with atomicAdd(&deg[u], 1)               (src/ours.cc:82-107)
//   repeat until every vertex is peeled; deg[] then holds the core numbers.
//
// Structure here (not the reference's): one persistent cooperative kernel runs every level.
//  * Frontier storage is one queue of V slots: each vertex is enqueued exactly once over the
//    whole peel, so the queue ends as the peel order and level starts are the k-shell starts.
//  * Select scans deg[] with 16-byte loads; matches are ranked by three 64-lane ballots and
//    one atomic per wave reserves queue space.
//  * Process: waves take 64-slot tickets from the queue head; a wave processes the vertices
//    of its ticket as their slots fill (cascades land in later tickets), flattening their
//    adjacency lists over the 64 lanes with a DPP prefix sum so short lists do not idle lanes.
//    Vertices with degree above a threshold are split into arc chunks in a second queue that
//    every wave drains first, so a hub is shared by the whole GPU.
//  * A level ends when the count of enqueued-but-unprocessed items reaches zero (items add
//    their children before retiring themselves, so zero is reached only when all work is done).
//  * Grid barriers separate select from process (a vertex must not be both selected by the
//    scan and enqueued by a decrement in the same level).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <new>
#include <vector>

#include "../../include/kcore.h"

#define AGENT __HIP_MEMORY_SCOPE_AGENT

namespace {

constexpr uint32_t kSent = 0xFFFFFFFFu;
constexpr uint64_t kSent64 = 0xFFFFFFFFFFFFFFFFull;
constexpr uint32_t kStride = 64;  // counters sit 256 bytes apart

// control block slots (index * kStride)
enum : uint32_t {
    C_TAIL = 0,      // queue slots reserved
    C_HEAD,          // queue tickets handed out
    C_OUT,           // items enqueued but not yet retired (vertices + arc chunks)
    C_BCNT,          // grid barrier arrivals
    C_BGEN,          // grid barrier generation
    C_ERR,           // 0 ok, 1 time limit, 2 no progress
    C_KMAX,
    C_LEVELS,
    C_SCANNED_LO,
    C_SCANNED_HI,
    C_CTAIL,         // arc-chunk slots reserved
    C_CHEAD,         // arc-chunk tickets handed out
    C_NEXTMIN,       // smallest remaining degree above the current level (level skipping)
    C_SNAP_NEXT0,    // next level, written by the last arriver of the level-end barrier,
    C_SNAP_NEXT1,    //   in a slot chosen by the parity of the level iteration
    C_SNAP_DONE0,    // 1 when every vertex is peeled (same parity scheme)
    C_SNAP_DONE1,
    C_COUNT
};

struct PeelArgs {
    const uint32_t* off;     // V+1
    const uint32_t* nbr;     // arcs
    const uint32_t* deg0;    // padded V, pad = kSent
    uint32_t* core;          // padded V, working degrees -> core numbers
    uint32_t* queue;         // V slots, kSent = empty
    uint64_t* chunks;        // chunk slots: (arc_begin | arc_end << 32), kSent64 = empty
    uint32_t* shell;         // per-level queue start, dmax + 2 entries
    uint32_t* ctrl;
    uint32_t V;
    uint32_t nvec;           // padded V / 4
    uint32_t dmax;
    uint32_t heavy;          // degree above which arcs are chunked
    uint32_t chunk_cap;      // chunk slots
    uint32_t flags;
    uint64_t time_limit;     // wall-clock ticks a single wait may take
};

__device__ __forceinline__ uint32_t lane_id() {
    return __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u));
}
// number of set bits of m below this lane
__device__ __forceinline__ uint32_t rank_below(uint64_t m) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}
__device__ __forceinline__ uint32_t bcast_first(uint32_t x) { return __builtin_amdgcn_readfirstlane(x); }

__device__ __forceinline__ uint32_t ld_acq(const uint32_t* p) {
    return __hip_atomic_load(p, __ATOMIC_ACQUIRE, AGENT);
}
__device__ __forceinline__ uint32_t ld_rlx(const uint32_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, AGENT);
}
__device__ __forceinline__ uint64_t ld_rlx64(const uint64_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, AGENT);
}

// Wave64 inclusive prefix sum with DPP: four row_shr steps give the scan inside each row of
// 16 lanes, row_bcast:15 and row_bcast:31 carry row totals into the following rows.
// Must be called with all 64 lanes active.
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t x) {
    int v = (int)x, t;
    t = __builtin_amdgcn_update_dpp(0, v, 0x111, 0xf, 0xf, false); v += t;  // row_shr:1
    t = __builtin_amdgcn_update_dpp(0, v, 0x112, 0xf, 0xf, false); v += t;  // row_shr:2
    t = __builtin_amdgcn_update_dpp(0, v, 0x114, 0xf, 0xf, false); v += t;  // row_shr:4
    t = __builtin_amdgcn_update_dpp(0, v, 0x118, 0xf, 0xf, false); v += t;  // row_shr:8
    t = __builtin_amdgcn_update_dpp(0, v, 0x142, 0xa, 0xf, false); v += t;  // row_bcast:15
    t = __builtin_amdgcn_update_dpp(0, v, 0x143, 0xc, 0xf, false); v += t;  // row_bcast:31
    return (uint32_t)v;
}

__device__ __forceinline__ uint32_t read_lane(uint32_t x, int lane) {
    return (uint32_t)__builtin_amdgcn_readlane((int)x, lane);
}

__device__ __forceinline__ void set_error(uint32_t* ctrl, uint32_t code) {
    __hip_atomic_fetch_max(ctrl + C_ERR * kStride, code, __ATOMIC_RELAXED, AGENT);
}
__device__ __forceinline__ bool has_error(const uint32_t* ctrl) {
    return ld_rlx(ctrl + C_ERR * kStride) != 0;
}

// Time-limited wait bookkeeping: every spin loop calls this; on expiry the error flag is set
// and all waits end, so the grid always drains.
struct Waiter {
    uint64_t t0 = 0;
    uint32_t n = 0;
    __device__ __forceinline__ bool expired(uint32_t* ctrl, uint64_t limit) {
        if ((++n & 63) != 0) return false;
        uint64_t now = wall_clock64();
        if (t0 == 0) t0 = now;
        if (now - t0 > limit) {
            set_error(ctrl, 1);
            return true;
        }
        return has_error(ctrl);
    }
};

// Generation barrier over all workgroups of the (co-resident) grid. The last block to arrive
// runs `hook` before releasing the others, so the hook sees every block's prior writes and
// its own writes are visible to every block after the barrier.
template <typename Hook>
__device__ void grid_barrier(uint32_t* ctrl, uint64_t limit, Hook hook) {
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t* cnt = ctrl + C_BCNT * kStride;
        uint32_t* gen = ctrl + C_BGEN * kStride;
        uint32_t g = ld_rlx(gen);
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        uint32_t a = __hip_atomic_fetch_add(cnt, 1u, __ATOMIC_ACQ_REL, AGENT);
        if (a == gridDim.x - 1) {
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
            hook();
            __hip_atomic_store(cnt, 0u, __ATOMIC_RELAXED, AGENT);
            __hip_atomic_store(gen, g + 1, __ATOMIC_RELEASE, AGENT);
        } else {
            Waiter w;
            while (ld_acq(gen) == g) {
                if (w.expired(ctrl, limit)) break;
                __builtin_amdgcn_s_sleep(1);
            }
        }
    }
    __syncthreads();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
}

// Reserve `cnt` consecutive slots of a queue for this wave; called by all lanes with the
// same cnt; returns the first slot.
__device__ __forceinline__ uint32_t wave_reserve(uint32_t* counter, uint32_t cnt) {
    uint32_t base = 0;
    if (lane_id() == 0) base = __hip_atomic_fetch_add(counter, cnt, __ATOMIC_RELAXED, AGENT);
    return bcast_first(base);
}

// ---------------------------------------------------------------------------------------
// select: append every v with core[v] == L (level 0 reads the input degrees and writes the
// working copy). Also records the smallest value above L for level skipping.
__device__ void select_level(const PeelArgs& a, uint32_t L, uint32_t* next_min_out) {
    const uint32_t lane = lane_id();
    const uint32_t gtid = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t gsize = gridDim.x * blockDim.x;
    uint32_t mn = kSent;
    const uint4* src = reinterpret_cast<const uint4*>(L == 0 ? a.deg0 : a.core);
    uint4* dst = reinterpret_cast<uint4*>(a.core);
    // every lane runs the same trip count so the ballots below see the whole wave
    const uint32_t trips = (a.nvec + gsize - 1) / gsize;
    for (uint32_t it = 0; it < trips; ++it) {
        const uint32_t i = it * gsize + gtid;
        uint4 d = make_uint4(kSent, kSent, kSent, kSent);
        if (i < a.nvec) {
            d = src[i];
            if (L == 0) dst[i] = d;
        }
        const uint32_t m0 = d.x == L, m1 = d.y == L, m2 = d.z == L, m3 = d.w == L;
        if (d.x > L && d.x < mn) mn = d.x;
        if (d.y > L && d.y < mn) mn = d.y;
        if (d.z > L && d.z < mn) mn = d.z;
        if (d.w > L && d.w < mn) mn = d.w;
        const uint32_t c = m0 + m1 + m2 + m3;
        const uint64_t b0 = __ballot(c & 1), b1 = __ballot(c & 2), b2 = __ballot(c & 4);
        if ((b0 | b1 | b2) == 0) continue;
        const uint32_t total = __popcll(b0) + 2 * __popcll(b1) + 4 * __popcll(b2);
        const uint32_t pre = rank_below(b0) + 2 * rank_below(b1) + 4 * rank_below(b2);
        uint32_t base = 0;
        if (lane == 0) {
            base = __hip_atomic_fetch_add(a.ctrl + C_TAIL * kStride, total, __ATOMIC_RELAXED, AGENT);
            __hip_atomic_fetch_add(a.ctrl + C_OUT * kStride, total, __ATOMIC_RELAXED, AGENT);
        }
        base = bcast_first(base) + pre;
        const uint32_t v0 = 4 * i;
        if (m0) a.queue[base++] = v0;
        if (m1) a.queue[base++] = v0 + 1;
        if (m2) a.queue[base++] = v0 + 2;
        if (m3) a.queue[base++] = v0 + 3;
    }
    *next_min_out = mn;
}

// Decrement the degrees of a set of arcs. Each active lane holds one neighbour u; crossing
// vertices are appended to the queue with one reservation per wave.
__device__ __forceinline__ void relax_and_append(const PeelArgs& a, uint32_t L, bool act,
                                                 uint32_t u, uint32_t& mn) {
    bool cross = false;
    if (act) {
        uint32_t du = a.core[u];  // may be stale-high, never stale-low: the atomic decides
        if (du > L) {
            uint32_t old = __hip_atomic_fetch_sub(a.core + u, 1u, __ATOMIC_RELAXED, AGENT);
            if (old == L + 1) {
                cross = true;
            } else if (old <= L) {
                __hip_atomic_fetch_add(a.core + u, 1u, __ATOMIC_RELAXED, AGENT);
            } else if (old - 1 < mn) {
                mn = old - 1;
            }
        }
    }
    const uint64_t m = __ballot(cross);
    if (m == 0) return;
    const uint32_t cnt = __popcll(m);
    uint32_t base = 0;
    if (lane_id() == 0) {
        __hip_atomic_fetch_add(a.ctrl + C_OUT * kStride, cnt, __ATOMIC_RELAXED, AGENT);
        base = __hip_atomic_fetch_add(a.ctrl + C_TAIL * kStride, cnt, __ATOMIC_RELAXED, AGENT);
    }
    base = bcast_first(base);
    // relaxed: the consumer needs only the id itself; barrier fences order everything else
    if (cross) __hip_atomic_store(a.queue + base + rank_below(m), u, __ATOMIC_RELAXED, AGENT);
}

// Process the vertices held by the lanes set in `ready` (lane's vertex v). Light vertices'
// arcs are flattened over the wave; heavy ones are cut into arc chunks for all waves.
// Returns the number of items this call enqueued as chunks (they retire separately).
__device__ void process_vertices(const PeelArgs& a, uint32_t L, uint64_t ready, uint32_t v,
                                 uint32_t* lds_incl, uint32_t* lds_start, uint32_t& mn) {
    const uint32_t lane = lane_id();
    const bool mine = (ready >> lane) & 1;
    uint32_t s = 0, d = 0;
    if (mine) {
        s = a.off[v];
        d = a.off[v + 1] - s;
    }
    // heavy vertices: push ceil(d / heavy) chunks
    const bool heavy = mine && d > a.heavy;
    const uint64_t hm = __ballot(heavy);
    if (hm) {
        uint32_t nch = heavy ? (d + a.heavy - 1) / a.heavy : 0;
        uint32_t incl = wave_incl_scan(nch);
        uint32_t tot = read_lane(incl, 63);
        uint32_t base = 0;
        if (lane == 0) {
            __hip_atomic_fetch_add(a.ctrl + C_OUT * kStride, tot, __ATOMIC_RELAXED, AGENT);
            base = __hip_atomic_fetch_add(a.ctrl + C_CTAIL * kStride, tot, __ATOMIC_RELAXED, AGENT);
        }
        base = bcast_first(base);
        if (heavy) {
            uint32_t slot = base + incl - nch;
            for (uint32_t k = 0; k < nch; ++k) {
                uint64_t jb = (uint64_t)s + (uint64_t)k * a.heavy;
                uint64_t je = jb + a.heavy;
                if (je > (uint64_t)s + d) je = (uint64_t)s + d;
                __hip_atomic_store(a.chunks + slot + k, jb | (je << 32), __ATOMIC_RELAXED, AGENT);
            }
            d = 0;
        }
    }
    const uint32_t incl = wave_incl_scan(d);
    const uint32_t T = read_lane(incl, 63);
    if (T == 0) return;
    lds_incl[lane] = incl;
    lds_start[lane] = s;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    for (uint32_t base = 0; base < T; base += 64) {
        const uint32_t j = base + lane;
        const bool act = j < T;
        uint32_t u = 0;
        if (act) {
            uint32_t k = 0;
#pragma unroll
            for (uint32_t step = 32; step; step >>= 1)
                if (lds_incl[k + step - 1] <= j) k += step;
            const uint32_t excl = k ? lds_incl[k - 1] : 0;
            u = a.nbr[lds_start[k] + (j - excl)];
        }
        relax_and_append(a, L, act, u, mn);
    }
    __builtin_amdgcn_wave_barrier();
}

// Process one arc chunk [jb, je) with the whole wave.
__device__ void process_chunk(const PeelArgs& a, uint32_t L, uint64_t ch, uint32_t& mn) {
    const uint32_t jb = (uint32_t)ch, je = (uint32_t)(ch >> 32);
    const uint32_t lane = lane_id();
    for (uint32_t base = jb; base < je; base += 64) {
        const uint32_t j = base + lane;
        const bool act = j < je;
        const uint32_t u = act ? a.nbr[j] : 0;
        relax_and_append(a, L, act, u, mn);
    }
}

__device__ __forceinline__ void retire(const PeelArgs& a, uint32_t n) {
    // same-address RMWs of one thread stay in program order, so an item's children are
    // counted before the item itself is retired
    if (lane_id() == 0) __hip_atomic_fetch_sub(a.ctrl + C_OUT * kStride, n, __ATOMIC_RELAXED, AGENT);
}

// Drain the current level: vertex tickets of 64 slots plus arc-chunk tickets of one slot.
__device__ void process_level(const PeelArgs& a, uint32_t L, uint32_t* lds_incl,
                              uint32_t* lds_start, uint32_t& mn) {
    const uint32_t lane = lane_id();
    uint32_t* tail = a.ctrl + C_TAIL * kStride;
    uint32_t* head = a.ctrl + C_HEAD * kStride;
    uint32_t* ctail = a.ctrl + C_CTAIL * kStride;
    uint32_t* chead = a.ctrl + C_CHEAD * kStride;
    uint32_t* out = a.ctrl + C_OUT * kStride;

    uint32_t b = wave_reserve(head, 64);
    uint64_t pending = 0;
    if (b < a.V) {
        uint32_t n = a.V - b;
        pending = n >= 64 ? ~0ull : ((1ull << n) - 1);
    }
    uint32_t ct = kSent;  // held chunk ticket
    Waiter w;
    while (true) {
        bool worked = false;
        // arc chunks first: they belong to hubs, the level's critical path
        if (ct == kSent) {
            uint32_t avail = 0;
            if (lane == 0) avail = ld_rlx(ctail) > ld_rlx(chead);
            if (bcast_first(avail)) ct = wave_reserve(chead, 1);
        }
        if (ct != kSent) {
            uint64_t ch = kSent64;
            if (ct < a.chunk_cap && lane == 0) ch = ld_rlx64(a.chunks + ct);
            uint32_t lo = bcast_first((uint32_t)ch), hi = bcast_first((uint32_t)(ch >> 32));
            ch = (uint64_t)lo | ((uint64_t)hi << 32);
            if (ch != kSent64) {
                process_chunk(a, L, ch, mn);
                retire(a, 1);
                ct = kSent;
                worked = true;
            }
        }
        if (pending) {
            uint32_t tl = 0;
            if (lane == 0) tl = ld_rlx(tail);
            tl = bcast_first(tl);
            uint32_t v = kSent;
            const uint32_t i = b + lane;
            if (((pending >> lane) & 1) && i < tl) v = ld_rlx(a.queue + i);
            const uint64_t ready = __ballot(v != kSent);
            if (ready) {
                process_vertices(a, L, ready, v, lds_incl, lds_start, mn);
                retire(a, __popcll(ready));
                pending &= ~ready;
                worked = true;
                if (!pending) {
                    b = wave_reserve(head, 64);
                    if (b < a.V) {
                        uint32_t n = a.V - b;
                        pending = n >= 64 ? ~0ull : ((1ull << n) - 1);
                    }
                }
            }
        }
        if (worked) {
            w = Waiter();
            continue;
        }
        uint32_t o = 0;
        if (lane == 0) o = ld_rlx(out);
        if (bcast_first(o) == 0) break;
        if (w.expired(a.ctrl, a.time_limit)) break;
        __builtin_amdgcn_s_sleep(1);
    }
}

// Level-end bookkeeping, run once per level by the last block to reach the level-end barrier:
// every block has finished processing, so tail and NEXTMIN are final for the level.
__device__ void level_end(const PeelArgs& a, uint32_t L, uint32_t par) {
    uint32_t* ctrl = a.ctrl;
    const uint32_t t = ld_acq(ctrl + C_TAIL * kStride);
    const uint32_t nm = ld_acq(ctrl + C_NEXTMIN * kStride);
    __hip_atomic_store(ctrl + C_NEXTMIN * kStride, kSent, __ATOMIC_RELAXED, AGENT);
    uint32_t done = t >= a.V;
    uint32_t next = L + 1;
    if ((a.flags & KC_FLAG_NO_LEVEL_SKIP) == 0 && nm != kSent && nm > next) next = nm;
    if (done) {
        a.shell[L + 1] = t;
    } else if (next > a.dmax) {
        set_error(ctrl, 2);
        done = 1;
    } else {
        for (uint32_t k = L + 1; k <= next; ++k) a.shell[k] = t;
        // tickets handed out beyond the filled slots are void: restart both queues at their tails
        __hip_atomic_store(ctrl + C_HEAD * kStride, t, __ATOMIC_RELAXED, AGENT);
        __hip_atomic_store(ctrl + C_CHEAD * kStride, ld_acq(ctrl + C_CTAIL * kStride),
                           __ATOMIC_RELAXED, AGENT);
    }
    __hip_atomic_store(ctrl + (C_SNAP_NEXT0 + par) * kStride, next, __ATOMIC_RELAXED, AGENT);
    __hip_atomic_store(ctrl + (C_SNAP_DONE0 + par) * kStride, done, __ATOMIC_RELAXED, AGENT);
}

template <int BLOCK>
__global__ __launch_bounds__(BLOCK) void peel_persistent(PeelArgs a) {
    __shared__ uint32_t s_incl[BLOCK / 64][64];
    __shared__ uint32_t s_start[BLOCK / 64][64];
    const uint32_t wid = threadIdx.x / 64;
    const bool leader = blockIdx.x == 0 && threadIdx.x == 0;
    uint32_t* ctrl = a.ctrl;
    if (leader) a.shell[0] = 0;

    uint32_t L = 0, levels = 0;
    for (uint32_t iter = 0;; ++iter) {
        uint32_t mn;
        select_level(a, L, &mn);
        levels++;
        grid_barrier(ctrl, a.time_limit, [] {});
        process_level(a, L, s_incl[wid], s_start[wid], mn);
        // smallest remaining degree above L, from the scan (untouched vertices) and from every
        // decrement result that stayed above L (touched vertices; their last result is lowest)
        for (int o = 32; o; o >>= 1) mn = min(mn, (uint32_t)__shfl_xor((int)mn, o));
        if (lane_id() == 0 && mn != kSent)
            __hip_atomic_fetch_min(ctrl + C_NEXTMIN * kStride, mn, __ATOMIC_RELAXED, AGENT);
        const uint32_t par = iter & 1;
        grid_barrier(ctrl, a.time_limit, [&] { level_end(a, L, par); });
        if (has_error(ctrl)) break;
        if (ld_acq(ctrl + (C_SNAP_DONE0 + par) * kStride)) break;
        L = ld_acq(ctrl + (C_SNAP_NEXT0 + par) * kStride);
    }
    if (leader) {
        const uint64_t scanned = (uint64_t)levels * a.nvec * 4ull;
        ctrl[C_KMAX * kStride] = L;
        ctrl[C_LEVELS * kStride] = levels;
        ctrl[C_SCANNED_LO * kStride] = (uint32_t)scanned;
        ctrl[C_SCANNED_HI * kStride] = (uint32_t)(scanned >> 32);
    }
}

template <typename T>
__global__ void fill_kernel(T* p, T v, uint64_t n) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    uint64_t s = (uint64_t)gridDim.x * blockDim.x;
    for (; i < n; i += s) p[i] = v;
}

__global__ void degrees_from_offsets(const uint32_t* off, uint32_t* deg, uint32_t V, uint32_t padV) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t s = gridDim.x * blockDim.x;
    for (; i < padV; i += s) deg[i] = i < V ? off[i + 1] - off[i] : kSent;
}

__global__ void selftest_scan_kernel(const uint32_t* in, uint32_t* out) {
    out[threadIdx.x] = wave_incl_scan(in[threadIdx.x]);
}

}  // namespace

// =======================================================================================
// host side

struct kc_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr, evk0 = nullptr, evk1 = nullptr;
    int cus = 0;
    uint64_t wall_rate_khz = 100000;
};

struct kc_graph {
    kc_ctx* ctx = nullptr;
    uint32_t V = 0, padV = 0, dmax = 0;
    uint64_t arcs = 0;
    uint32_t* d_off = nullptr;
    uint32_t* d_nbr = nullptr;
    uint32_t* d_deg0 = nullptr;
    uint32_t* d_core = nullptr;
    uint32_t* d_queue = nullptr;
    uint64_t* d_chunks = nullptr;
    uint32_t chunk_cap = 0;
    uint32_t* d_shell = nullptr;
    uint32_t* d_ctrl = nullptr;
    uint32_t last_kmax = 0;
    bool peeled = false;
};

static kc_status hip_status(hipError_t e) {
    if (e == hipSuccess) return KC_OK;
    if (e == hipErrorOutOfMemory) return KC_ERR_OOM;
    if (e == hipErrorNoDevice || e == hipErrorInvalidDevice) return KC_ERR_NO_DEVICE;
    return KC_ERR_HIP;
}
#define KC_TRY(x)                                 \
    do {                                          \
        hipError_t e_ = (x);                      \
        if (e_ != hipSuccess) return hip_status(e_); \
    } while (0)

template <int BLOCK>
static kc_status launch_peel(kc_ctx* c, const PeelArgs& args, uint32_t* grid_out) {
    int per_cu = 0;
    KC_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, peel_persistent<BLOCK>, BLOCK, 0));
    if (per_cu < 1) return KC_ERR_HIP;
    uint32_t grid = (uint32_t)per_cu * (uint32_t)c->cus;
    // keep head tickets inside uint32: within a level head <= tail + 64 * (waves + V / 64 + 1)
    if (2ull * args.V + 128ull * grid * (BLOCK / 64) + 64 >= 0xFFFFFFFFull) return KC_ERR_CAPACITY;
    *grid_out = grid;
    PeelArgs a = args;
    void* kargs[] = {&a};
    KC_TRY(hipLaunchCooperativeKernel(reinterpret_cast<void*>(peel_persistent<BLOCK>), dim3(grid),
                                      dim3(BLOCK), kargs, 0, c->stream));
    return KC_OK;
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
    int rate = 0;
    if (hipDeviceGetAttribute(&rate, hipDeviceAttributeWallClockRate, device) == hipSuccess && rate > 0)
        c->wall_rate_khz = (uint64_t)rate;
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
    if (c->ev0) (void)hipEventDestroy(c->ev0);
    if (c->ev1) (void)hipEventDestroy(c->ev1);
    if (c->evk0) (void)hipEventDestroy(c->evk0);
    if (c->evk1) (void)hipEventDestroy(c->evk1);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

void kc_graph_free(kc_graph* g) {
    if (!g) return;
    (void)hipSetDevice(g->ctx->device);
    (void)hipStreamSynchronize(g->ctx->stream);
    (void)hipFree(g->d_off);
    (void)hipFree(g->d_nbr);
    (void)hipFree(g->d_deg0);
    (void)hipFree(g->d_core);
    (void)hipFree(g->d_queue);
    (void)hipFree(g->d_chunks);
    (void)hipFree(g->d_shell);
    (void)hipFree(g->d_ctrl);
    delete g;
}

kc_status kc_graph_upload(kc_ctx* c, const kc_csr* h, kc_graph** out) {
    if (!c || !h || !out || !h->offsets || (h->arcs && !h->neighbors)) return KC_ERR_ARG;
    if (h->V == 0 || h->V >= 0x7FFFFFFFu || h->arcs >= 0xFFFFFFFFull) return KC_ERR_CAPACITY;
    if (h->offsets[0] != 0 || h->offsets[h->V] != h->arcs) return KC_ERR_ARG;
    KC_TRY(hipSetDevice(c->device));
    kc_graph* g = new (std::nothrow) kc_graph;
    if (!g) return KC_ERR_OOM;
    g->ctx = c;
    g->V = h->V;
    g->arcs = h->arcs;
    g->padV = (h->V + 255u) & ~255u;
    uint32_t dmax = 0;
    for (uint32_t v = 0; v < h->V; ++v) dmax = std::max(dmax, h->offsets[v + 1] - h->offsets[v]);
    g->dmax = dmax;
    // arc chunks: only vertices with degree d > heavy (>= 256) are cut, into ceil(d / heavy)
    // < 2 d / heavy chunks each, and each vertex is cut at most once per peel
    g->chunk_cap = (uint32_t)std::min<uint64_t>(2 * (h->arcs / 256) + 64, 0xFFFFFFF0ull);
    auto fail = [&](hipError_t e) {
        kc_graph_free(g);
        return hip_status(e);
    };
    hipError_t e;
    if ((e = hipMalloc(&g->d_off, ((size_t)h->V + 1) * 4)) != hipSuccess) return fail(e);
    if ((e = hipMalloc(&g->d_nbr, std::max<size_t>(h->arcs, 1) * 4)) != hipSuccess) return fail(e);
    if ((e = hipMalloc(&g->d_deg0, (size_t)g->padV * 4)) != hipSuccess) return fail(e);
    if ((e = hipMalloc(&g->d_core, (size_t)g->padV * 4)) != hipSuccess) return fail(e);
    if ((e = hipMalloc(&g->d_queue, (size_t)g->padV * 4)) != hipSuccess) return fail(e);
    if ((e = hipMalloc(&g->d_chunks, (size_t)g->chunk_cap * 8)) != hipSuccess) return fail(e);
    if ((e = hipMalloc(&g->d_shell, ((size_t)dmax + 3) * 4)) != hipSuccess) return fail(e);
    if ((e = hipMalloc(&g->d_ctrl, (size_t)C_COUNT * kStride * 4)) != hipSuccess) return fail(e);
    if ((e = hipMemcpyAsync(g->d_off, h->offsets, ((size_t)h->V + 1) * 4, hipMemcpyHostToDevice, c->stream)) != hipSuccess) return fail(e);
    if (h->arcs && (e = hipMemcpyAsync(g->d_nbr, h->neighbors, h->arcs * 4, hipMemcpyHostToDevice, c->stream)) != hipSuccess) return fail(e);
    if (h->degrees) {
        if ((e = hipMemcpyAsync(g->d_deg0, h->degrees, (size_t)h->V * 4, hipMemcpyHostToDevice, c->stream)) != hipSuccess) return fail(e);
        if (g->padV > h->V) {
            fill_kernel<uint32_t><<<1, 256, 0, c->stream>>>(g->d_deg0 + h->V, kSent, g->padV - h->V);
        }
    } else {
        degrees_from_offsets<<<1024, 256, 0, c->stream>>>(g->d_off, g->d_deg0, g->V, g->padV);
    }
    if ((e = hipGetLastError()) != hipSuccess) return fail(e);
    if ((e = hipStreamSynchronize(c->stream)) != hipSuccess) return fail(e);
    *out = g;
    return KC_OK;
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
    uint32_t block = prm.block ? prm.block : 256;
    if (block != 256 && block != 512 && block != 1024) return KC_ERR_ARG;
    uint32_t heavy = prm.heavy_threshold ? prm.heavy_threshold : 4096;
    if (heavy < 256) heavy = 256;

    PeelArgs a{};
    a.off = g->d_off;
    a.nbr = g->d_nbr;
    a.deg0 = g->d_deg0;
    a.core = g->d_core;
    a.queue = g->d_queue;
    a.chunks = g->d_chunks;
    a.shell = g->d_shell;
    a.ctrl = g->d_ctrl;
    a.V = g->V;
    a.nvec = g->padV / 4;
    a.dmax = g->dmax;
    a.heavy = heavy;
    a.chunk_cap = g->chunk_cap;
    a.flags = prm.flags;
    a.time_limit = c->wall_rate_khz * 1000ull * 20;  // 20 s per wait

    KC_TRY(hipEventRecord(c->ev0, c->stream));
    // reset state: queue and chunk slots empty, counters zero, NEXTMIN = none
    fill_kernel<uint32_t><<<2048, 256, 0, c->stream>>>(g->d_queue, kSent, g->padV);
    fill_kernel<uint64_t><<<2048, 256, 0, c->stream>>>(g->d_chunks, kSent64, g->chunk_cap);
    fill_kernel<uint32_t><<<8, 256, 0, c->stream>>>(g->d_ctrl, 0u, (uint64_t)C_COUNT * kStride);
    fill_kernel<uint32_t><<<1, 64, 0, c->stream>>>(g->d_ctrl + C_NEXTMIN * kStride, kSent, 1);
    fill_kernel<uint32_t><<<1, 64, 0, c->stream>>>(g->d_shell, 0u, (uint64_t)g->dmax + 3);
    KC_TRY(hipGetLastError());
    uint32_t grid = 0;
    kc_status s;
    KC_TRY(hipEventRecord(c->evk0, c->stream));
    switch (block) {
        case 256: s = launch_peel<256>(c, a, &grid); break;
        case 512: s = launch_peel<512>(c, a, &grid); break;
        default: s = launch_peel<1024>(c, a, &grid); break;
    }
    if (s != KC_OK) return s;
    KC_TRY(hipEventRecord(c->evk1, c->stream));
    KC_TRY(hipEventRecord(c->ev1, c->stream));
    uint32_t ctrl_h[C_COUNT * kStride];
    KC_TRY(hipMemcpyAsync(ctrl_h, g->d_ctrl, sizeof(ctrl_h), hipMemcpyDeviceToHost, c->stream));
    KC_TRY(hipStreamSynchronize(c->stream));
    float ms = 0, msk = 0;
    KC_TRY(hipEventElapsedTime(&ms, c->ev0, c->ev1));
    KC_TRY(hipEventElapsedTime(&msk, c->evk0, c->evk1));
    uint32_t err = ctrl_h[C_ERR * kStride];
    if (err == 1) return KC_ERR_DEVICE;
    if (err == 2) return KC_ERR_NO_PROGRESS;
    g->last_kmax = ctrl_h[C_KMAX * kStride];
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
        st->levels = ctrl_h[C_LEVELS * kStride];
        uint64_t scanned = (uint64_t)ctrl_h[C_SCANNED_LO * kStride] |
                           ((uint64_t)ctrl_h[C_SCANNED_HI * kStride] << 32);
        st->scanned_entries = scanned;
        st->bytes_scan = 4 * scanned;
        st->bytes_edge = 16ull * g->V + 20ull * (g->arcs / 2);
        st->bytes_alg = st->bytes_scan + st->bytes_edge;
        st->grid = grid;
        st->block = block;
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
        KC_TRY(hipMemcpyAsync(order_out, g->d_queue, (size_t)g->V * 4, hipMemcpyDeviceToHost, c->stream));
    if (shell_off_out)
        KC_TRY(hipMemcpyAsync(shell_off_out, g->d_shell, ((size_t)g->last_kmax + 2) * 4,
                              hipMemcpyDeviceToHost, c->stream));
    KC_TRY(hipStreamSynchronize(c->stream));
    return KC_OK;
}

const uint32_t* kc_graph_core_device(const kc_graph* g) { return g ? g->d_core : nullptr; }

// Checks the wave64 DPP prefix sum against a host prefix sum. Returns KC_OK when equal.
kc_status kc_selftest(kc_ctx* c) {
    if (!c) return KC_ERR_ARG;
    KC_TRY(hipSetDevice(c->device));
    uint32_t h_in[64], h_out[64];
    for (int i = 0; i < 64; ++i) h_in[i] = (uint32_t)(i * 7 + 3) % 11 + (i == 40 ? 100000u : 0u);
    uint32_t *d_in = nullptr, *d_out = nullptr;
    KC_TRY(hipMalloc(&d_in, 256));
    KC_TRY(hipMalloc(&d_out, 256));
    (void)hipMemcpy(d_in, h_in, 256, hipMemcpyHostToDevice);
    selftest_scan_kernel<<<1, 64, 0, c->stream>>>(d_in, d_out);
    (void)hipStreamSynchronize(c->stream);
    (void)hipMemcpy(h_out, d_out, 256, hipMemcpyDeviceToHost);
    (void)hipFree(d_in);
    (void)hipFree(d_out);
    uint32_t acc = 0;
    for (int i = 0; i < 64; ++i) {
        acc += h_in[i];
        if (h_out[i] != acc) return KC_ERR_DEVICE;
    }
    return KC_OK;
}

}  // extern "C"
