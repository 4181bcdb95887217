// peel_kernel.h — the single-graph persistent peeling kernel for gfx950.
//
// Method: the reference's level-synchronous peel (src/ours.cc:2-160). For level L:
//   select  every v with deg[v] == L joins the frontier             (src/ours.cc:2-33)
//   process for each frontier v and neighbour u with deg[u] > L:
//           a = atomicSub(&deg[u], 1); a == L+1 -> u joins the frontier of this level;
//           a <= L -> undo with atomicAdd(&deg[u], 1)                (src/ours.cc:82-107)
//   until every vertex is peeled; deg[] then holds the core numbers.
//
// Structure (this engine's own): one cooperative launch runs every level.
//  * Global hot spots are what limits this workload on MI355X: one memory address touched by
//    thousands of waves serialises at the memory side. So the kernel touches shared counters
//    once per 64 items at most and partitions work statically wherever it can.
//  * select: each wave streams its share of deg[] with four 16-byte loads in flight per lane;
//    matches are staged in the wave's LDS stage and flushed 64 at a time into order[] (the peel
//    order) with one reservation; heavy ones (degree above `heavy`) are also cut into arc
//    chunks.
//  * process, in sub-rounds separated by grid barriers. A sub-round's work is fixed when it
//    starts: a vertex range (the level's order[] entries after select, later the spill queue
//    work[]) and a chunk range, both dealt round-robin over every wave of the grid in batches
//    of 64 vertices (adjacency flattened over the lanes with a DPP prefix sum) or one chunk.
//    A vertex that crosses to L is staged in LDS and flushed 64 at a time: recorded in
//    order[], then cut into chunks (heavy), kept in the workgroup's LDS queue for its next
//    local round (light, within an arc budget and a round limit), or spilled to work[] for the
//    next sub-round. A level ends at the first sub-round that leaves no spill and no chunk.
//  * grid barrier: two-level arrival (one counter per XCD group, then one for the grid), the
//    last workgroup runs the level bookkeeping with all of its threads and publishes the next
//    sub-round in a parity-indexed snapshot, then raises one release flag per XCD group;
//    waiters poll their group's flag with relaxed loads. Counters sit 8 KB apart.
//  * coherence without fences: an agent-scope fence writes back or invalidates a whole L2 and
//    costs ~70 us per barrier on MI355X. Every datum one workgroup writes and another reads
//    (queue slots, chunks, snapshots, per-workgroup minima) is accessed with device-scope
//    (SC1) loads and stores, degrees change only by device-scope atomics and are scanned with
//    SC1 buffer loads, and a barrier only waits for the wave's own memory operations to
//    complete. Read-only arrays (offsets, adjacency, input degrees) stay cacheable.
//  * elastic start: workgroups that are not resident when registration closes exit, and the
//    rest index themselves by check-in rank (see register_workgroup).
//  * level skipping: the next level is the smallest degree left above L, from the scan
//    (untouched vertices) and the decrement results (touched vertices), so levels with an
//    empty frontier, which the reference still scans, are not scanned.
#pragma once
#include "device_common.h"

namespace kcdev {

constexpr uint32_t kStride = 2048;  // control words sit 8 KB apart
constexpr uint32_t kXcds = 8;       // MI355X: workgroups are dealt round-robin to 8 XCDs
constexpr uint32_t kStageCap = 256; // per-wave staging entries

enum : uint32_t {
    C_OTAIL = 0,  // order[] records appended
    C_WTAIL,      // work[] spills appended
    C_CTAIL,      // chunks appended
    C_ERR,        // 0 ok, 1 time limit, 2 no progress
    C_LSTART,     // order[] position where the current level starts
    C_WSTART,     // work[] position where the current sub-round's spills start
    C_CSTART,     // chunk position where the current sub-round's new chunks start
    C_BTOP,       // top-level barrier arrivals (one per XCD group)
    C_LEVELS,
    C_SUBROUNDS,
    C_NTRACE,
    C_CHECKIN,    // workgroups checked in at kernel start (bit 31 = registration closed)
    C_NPART,      // participating workgroups, published when registration closes
    C_RESULT,     // summary written at exit: see R_* below
    C_BXCD,       // kXcds per-XCD-group arrival counters
    C_REL = C_BXCD + kXcds,  // kXcds per-XCD-group release flags (barrier generation)
    C_SNAP = C_REL + kXcds,  // two snapshot slots
    C_COUNT = C_SNAP + 2
};
// Result words. From R_WORDS on: dense entries scanned, list entries scanned and alive-list
// entries written (cascade-queue kernels), chunks queued (every kernel), and claims of the next
// level held for it (level hand-off only, not read by the host).
enum : uint32_t { R_ERR = 0, R_KMAX, R_LEVELS, R_SUBROUNDS, R_NPART, R_NTRACE, R_WORDS };

enum : uint32_t { PH_SUBROUND = 1, PH_NEXT_LEVEL = 2, PH_DONE = 3 };
enum : uint32_t { S_PHASE = 0, S_SRC, S_VBEG, S_VEND, S_CBEG, S_CEND, S_NEXT, S_WORDS };
enum : uint32_t { TR_SCAN = 1, TR_SUB = 2, TR_LEVEL = 3 };

struct PeelArgs {
    const uint32_t* off;   // V+1
    const uint32_t* nbr;   // arcs
    const uint32_t* deg0;  // padded V, pad = kSent
    uint32_t* core;        // padded V, working degrees -> core numbers
    uint32_t* order;       // V: peel order
    uint32_t* work;        // V: spilled frontier vertices
    uint64_t* chunks;      // arc chunks (begin | end << 32)
    uint32_t* shell;       // dmax + 3: order[] start of each level
    uint32_t* wgmin;       // per-workgroup smallest degree above the level
    uint64_t* trace;       // (kind << 56 | subround << 32 | level, wall clock) pairs
    uint32_t* dbg;         // 8 words per workgroup: loop position, for diagnosing a time-out
    uint32_t* ctrl;
    uint32_t V;
    uint32_t nvec;         // padded V / 4
    uint32_t dmax;
    uint32_t heavy;        // degree above which a vertex's arcs are processed as chunks
    uint32_t chunk;        // arcs per chunk
    uint32_t chunk_cap;
    uint32_t lq_depth;     // LDS queue entries per local round (0 = no local rounds)
    uint32_t lq_arcs;      // arcs a workgroup may keep locally per local round
    uint32_t lq_rounds;    // local rounds per sub-round
    uint32_t flags;
    uint32_t trace_cap;
    uint64_t time_limit;   // wall-clock ticks one wait may take
    uint64_t reg_window;   // wall-clock ticks without a new check-in that close registration
    // cascade-queue kernel (peel_async) only
    uint32_t* alive[2];    // padded V each: lists of vertices still unpeeled, for compacted scans
    uint32_t async_mode;   // 1: crossings are consumed from order[] itself (no work[] spill)
    uint32_t compact_ratio;  // compact when unpeeled * ratio < entries the scan reads
    uint32_t poll_max;     // longest sleep between polls of an idle workgroup (0..3)
};

constexpr uint32_t FLAG_NO_LEVEL_SKIP = 1u;
constexpr uint32_t FLAG_TRACE = 1u << 8;
constexpr uint32_t FLAG_DIRECT_POLL = 1u << 10;

__device__ __forceinline__ uint32_t* cw(const PeelArgs& a, uint32_t slot) { return a.ctrl + slot * kStride; }

// Per-workgroup LDS state.
template <int BLOCK>
struct WgState {
    static constexpr int WPB = BLOCK / 64;
    static constexpr int LQ = BLOCK * 2;  // local queue entries per buffer
    uint32_t stage[WPB][kStageCap];
    uint32_t lq[2][LQ];
    uint32_t incl[WPB][64];
    uint32_t start[WPB][64];
    uint32_t cur_n, head, nxt_n, cur, larcs;
    uint32_t red[WPB];
    uint32_t is_last;
    uint32_t mn;     // smallest degree above L seen by this workgroup in the current level
    uint32_t gen;    // barrier generation read on arrival
    uint32_t vbid;   // this workgroup's index among the participants
    uint32_t npart;  // participating workgroups
    uint32_t appended;  // cascade-queue kernel: a wave appended crossings since the last hint
};

constexpr uint32_t kClosed = 0x80000000u;

// Elastic start. A cooperative launch is accepted by the runtime even when fewer workgroups
// than requested can be resident at once, and a grid barrier would then wait forever for
// workgroups that can only start after others exit. So workgroups check in first; the first
// one closes registration once every workgroup has checked in, or once no new one has for
// `reg_window` ticks, and publishes the participant count. Workgroups that check in after the
// close exit at once; the rest use their check-in rank as index and the participant count as
// grid size. Returns false for a workgroup that must exit.
template <int BLOCK>
__device__ bool register_workgroup(const PeelArgs& a, WgState<BLOCK>& S) {
    if (threadIdx.x == 0) {
        uint32_t* ci = cw(a, C_CHECKIN);
        const uint32_t old = add_rlx(ci, 1u);
        S.vbid = kSent;
        S.npart = 0;
        if (!(old & kClosed)) {
            S.vbid = old;
            if (old == 0) {
                uint32_t seen = 1;
                uint64_t t_last = wall_clock64();
                while (true) {
                    const uint32_t c = ld_rlx(ci) & ~kClosed;
                    const uint64_t now = wall_clock64();
                    if (c != seen) {
                        seen = c;
                        t_last = now;
                    }
                    if (c >= gridDim.x || now - t_last > a.reg_window) break;
                    __builtin_amdgcn_s_sleep(2);
                }
                const uint32_t p = __hip_atomic_fetch_or(ci, kClosed, __ATOMIC_ACQ_REL, KC_AGENT) & ~kClosed;
                __hip_atomic_store(cw(a, C_NPART), p, __ATOMIC_RELEASE, KC_AGENT);
            }
            Waiter w;
            uint32_t p;
            while ((p = ld_rlx(cw(a, C_NPART))) == 0) {
                if (w.expired(cw(a, C_ERR), a.time_limit)) break;
                __builtin_amdgcn_s_sleep(2);
            }
            S.npart = p;
        }
    }
    __syncthreads();
    return S.vbid != kSent && S.vbid < S.npart;
}

// ---------------------------------------------------------------------------------------
// Grid barrier with a hook run by every thread of the last workgroup to arrive.
template <int BLOCK, typename Hook>
__device__ void grid_barrier(const PeelArgs& a, WgState<BLOCK>& S, Hook hook) {
    drain_memory();  // this wave's SC1 stores and atomics are complete, hence visible
    __syncthreads();
    if (threadIdx.x == 0) {
        const uint32_t P = S.npart;
        const uint32_t x = S.vbid % kXcds;
        const uint32_t g = ld_rlx(cw(a, C_REL + x));  // generation before arriving
        const uint32_t members = (P - x + kXcds - 1) / kXcds;  // participants in this group
        const uint32_t groups = P < kXcds ? P : kXcds;
        uint32_t last = 0;
        const uint32_t ax = add_rlx(cw(a, C_BXCD + x), 1u);
        if (ax == members - 1) {
            st_rlx(cw(a, C_BXCD + x), 0u);
            drain_memory();
            if (add_rlx(cw(a, C_BTOP), 1u) == groups - 1) {
                st_rlx(cw(a, C_BTOP), 0u);
                last = 1;
            }
        }
        st_rlx(a.dbg + S.vbid * 8 + 4, ax);
        S.is_last = last;
        S.gen = g;
    }
    __syncthreads();
    if (S.is_last) {
        hook();
        drain_memory();
        __syncthreads();
        if (threadIdx.x == 0) {
            for (uint32_t x = 0; x < kXcds; ++x) st_rlx(cw(a, C_REL + x), S.gen + 1);
        }
    } else if (threadIdx.x == 0) {
        uint32_t* rel = cw(a, C_REL + S.vbid % kXcds);
        const uint32_t g = S.gen;
        Waiter w;
        while (ld_rlx(rel) == g) {
            if (w.expired(cw(a, C_ERR), a.time_limit)) {
                st_rlx(a.dbg + S.vbid * 8 + 7, 0xDEADu);
                break;
            }
            __builtin_amdgcn_s_sleep(2);
        }
    }
    __syncthreads();
}

__device__ __forceinline__ void trace_rec(const PeelArgs& a, uint32_t kind, uint32_t sub, uint32_t level) {
    if (!(a.flags & FLAG_TRACE)) return;
    uint32_t i = add_rlx(cw(a, C_NTRACE), 1u);
    if (i < a.trace_cap) {
        a.trace[2 * i] = ((uint64_t)kind << 56) | ((uint64_t)sub << 32) | level;
        a.trace[2 * i + 1] = wall_clock64();
    }
}

// ---------------------------------------------------------------------------------------
// Arc chunks of heavy vertices. Lanes with `heavy` set push ceil(d / a.chunk) chunks of the
// arc range [s, s + d) with one reservation for the wave. All lanes call.
__device__ __forceinline__ void push_chunks(const PeelArgs& a, bool heavy, uint32_t s, uint32_t d) {
    if (__ballot(heavy) == 0) return;
    const uint32_t nch = heavy ? (d + a.chunk - 1) / a.chunk : 0;
    const uint32_t incl = wave_incl_scan(nch);
    const uint32_t base = wave_reserve(cw(a, C_CTAIL), read_lane(incl, 63));
    if (heavy) {
        const uint32_t slot = base + incl - nch;
        for (uint32_t k = 0; k < nch; ++k) {
            const uint64_t jb = (uint64_t)s + (uint64_t)k * a.chunk;
            uint64_t je = jb + a.chunk;
            if (je > (uint64_t)s + d) je = (uint64_t)s + d;
            st_rlx64(a.chunks + slot + k, jb | (je << 32));
        }
    }
}

// Wave-aggregated append of the lanes' v (where pred) to a global queue.
__device__ __forceinline__ void append_global(uint32_t* q, uint32_t* counter, bool pred, uint32_t v) {
    const uint64_t m = __ballot(pred);
    if (m == 0) return;
    const uint32_t base = wave_reserve(counter, __popcll(m));
    if (pred) st_rlx(q + base + rank_below(m), v);
}

// Flush the wave's stage of n vertices. Every one is recorded in order[]; heavy ones become
// chunks; for crossings (not select matches) light ones go to the workgroup's next local round
// when `local` allows and the arc budget has room, else to work[]. All lanes call; n uniform.
template <int BLOCK>
__device__ void flush_stage(const PeelArgs& a, WgState<BLOCK>& S, uint32_t& n, bool crossings, bool local) {
    if (n == 0) return;
    const uint32_t lane = lane_id();
    const uint32_t* st = S.stage[threadIdx.x / 64];
    const uint32_t rb = wave_reserve(cw(a, C_OTAIL), n);
    for (uint32_t i = lane; i < n; i += 64) st_rlx(a.order + rb + i, st[i]);
    if (a.async_mode && crossings && lane == 0) S.appended = 1;
    for (uint32_t i0 = 0; i0 < n; i0 += 64) {
        const bool has = i0 + lane < n;
        const uint32_t u = has ? st[i0 + lane] : 0;
        const uint32_t d = has ? a.deg0[u] : 0;  // original degree = adjacency length
        const bool heavy = has && d > a.heavy;
        if (__ballot(heavy)) push_chunks(a, heavy, heavy ? a.off[u] : 0, d);
        if (!crossings || a.async_mode) continue;
        const bool light = has && !heavy;
        bool spill = light;
        if (local && __ballot(light)) {
            const uint32_t incl = wave_incl_scan(light ? d : 0);
            uint32_t abase = 0;
            const uint32_t wave_arcs = read_lane(incl, 63);
            if (lane == 0) abase = __hip_atomic_fetch_add(&S.larcs, wave_arcs, __ATOMIC_RELAXED, KC_WG);
            abase = bcast_first(abase);
            const bool fits = light && abase + incl <= a.lq_arcs;
            const uint64_t fm = __ballot(fits);
            if (fm) {
                uint32_t base = 0;
                if (lane == 0) base = __hip_atomic_fetch_add(&S.nxt_n, (uint32_t)__popcll(fm), __ATOMIC_RELAXED, KC_WG);
                base = bcast_first(base);
                const uint32_t slot = base + rank_below(fm);
                if (fits && slot < a.lq_depth) {
                    S.lq[S.cur ^ 1][slot] = u;
                    spill = false;
                }
            }
        }
        append_global(a.work, cw(a, C_WTAIL), spill, u);
    }
    n = 0;
    __builtin_amdgcn_wave_barrier();
}

// Stage the lanes' u where pred (all lanes call; n is the wave's stage fill, uniform).
template <int BLOCK>
__device__ __forceinline__ void stage_push(const PeelArgs& a, WgState<BLOCK>& S, uint32_t& n, bool pred,
                                           uint32_t u, bool crossings, bool local) {
    const uint64_t m = __ballot(pred);
    if (m == 0) return;
    const uint32_t cnt = __popcll(m);
    if (n + cnt > kStageCap) flush_stage(a, S, n, crossings, local);
    if (pred) S.stage[threadIdx.x / 64][n + rank_below(m)] = u;
    n += cnt;
    __builtin_amdgcn_wave_barrier();
}

constexpr uint32_t kU = 4;  // arcs per lane in flight per iteration

// kU*64 arcs of the current level, one group of kU per lane: all adjacency and degree loads
// are issued before the decrements. Per arc: precheck (a stale read is only ever too high, so
// the atomic decides), decrement, undo when it went below the level. Crossing arcs are staged.
template <int BLOCK>
__device__ __forceinline__ void relax_group(const PeelArgs& a, WgState<BLOCK>& S, uint32_t L,
                                            const bool (&act)[kU], const uint32_t (&u)[kU],
                                            uint32_t& mn, uint32_t& sn, bool local) {
    // The kU prechecks are loaded unconditionally, so they leave back to back and each
    // decrement waits only for its own load. (Under a per-arc branch the compiler waits for
    // each load before issuing the next.) Inactive lanes read the graph's first offset, a
    // read-only word, instead of a degree that decrements may be contending for.
    uint32_t du[kU];
#pragma unroll
    for (uint32_t k = 0; k < kU; ++k) {
        const uint32_t d = *(act[k] ? a.core + u[k] : a.off);
        du[k] = act[k] ? d : 0;
    }
    uint32_t old[kU];
#pragma unroll
    for (uint32_t k = 0; k < kU; ++k)
        old[k] = du[k] > L ? __hip_atomic_fetch_sub(a.core + u[k], 1u, __ATOMIC_RELAXED, KC_AGENT) : 0;
#pragma unroll
    for (uint32_t k = 0; k < kU; ++k) {
        const bool cross = du[k] > L && old[k] == L + 1;
        if (du[k] > L && old[k] <= L) {
            __hip_atomic_fetch_add(a.core + u[k], 1u, __ATOMIC_RELAXED, KC_AGENT);
        } else if (du[k] > L && old[k] > L + 1 && old[k] - 1 < mn) {
            mn = old[k] - 1;
        }
        stage_push(a, S, sn, cross, u[k], true, local);
    }
}

// Process up to 64 vertices, one per lane where `mine`. Heavy vertices are skipped (their
// chunks were queued when they joined the frontier). Arcs are flattened over the wave.
template <int BLOCK, bool PREFETCH>
__device__ void process_vertices(const PeelArgs& a, WgState<BLOCK>& S, uint32_t L, bool mine, uint32_t v,
                                 uint32_t& mn, uint32_t& sn, bool local) {
    const uint32_t lane = lane_id();
    const uint32_t wid = threadIdx.x / 64;
    uint32_t s = 0, d = 0;
    if (mine) {
        s = a.off[v];
        d = a.off[v + 1] - s;
        if (d > a.heavy) d = 0;
    }
    const uint32_t incl = wave_incl_scan(d);
    const uint32_t T = read_lane(incl, 63);
    if (T == 0) return;
    uint32_t* lin = S.incl[wid];
    uint32_t* lst = S.start[wid];
    lin[lane] = incl;
    lst[lane] = s;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    // The neighbour at flattened positions j[0..kU): kU binary searches over the wave's
    // inclusive prefix sums, advanced step by step together so each step's kU LDS reads are
    // independent, then the kU adjacency loads.
    auto arcs = [&](const uint32_t (&j)[kU], uint32_t (&out)[kU]) {
        uint32_t at[kU];
#pragma unroll
        for (uint32_t k = 0; k < kU; ++k) at[k] = 0;
#pragma unroll
        for (uint32_t step = 32; step; step >>= 1) {
#pragma unroll
            for (uint32_t k = 0; k < kU; ++k)
                if (lin[at[k] + step - 1] <= j[k]) at[k] += step;
        }
#pragma unroll
        for (uint32_t k = 0; k < kU; ++k) {
            const uint32_t excl = at[k] ? lin[at[k] - 1] : 0;
            out[k] = a.nbr[lst[at[k]] + (j[k] - excl)];
        }
    };
    // Adjacency loads are unconditional too (inactive lanes read arc 0, which exists since
    // T > 0), so a group's kU loads leave back to back; whether a next group exists is the
    // same for the whole wave.
    bool act[kU];
    uint32_t u[kU], j[kU];
#pragma unroll
    for (uint32_t k = 0; k < kU; ++k) {
        act[k] = k * 64 + lane < T;
        j[k] = act[k] ? k * 64 + lane : 0;
    }
    arcs(j, u);
    for (uint32_t base = 0; base < T; base += 64 * kU) {
        const bool more = base + 64 * kU < T;
        bool actn[kU];
        uint32_t un[kU];
#pragma unroll
        for (uint32_t k = 0; k < kU; ++k) {
            const uint32_t jn = base + 64 * kU + k * 64 + lane;
            actn[k] = jn < T;
            j[k] = actn[k] ? jn : 0;
            un[k] = 0;
        }
        if (PREFETCH && more) arcs(j, un);  // next group's loads fly during this one's atomics
        relax_group(a, S, L, act, u, mn, sn, local);
        if (!PREFETCH && more) arcs(j, un);
#pragma unroll
        for (uint32_t k = 0; k < kU; ++k) {
            u[k] = un[k];
            act[k] = actn[k];
        }
    }
    __builtin_amdgcn_wave_barrier();
}

template <int BLOCK, bool PREFETCH>
__device__ void process_chunk(const PeelArgs& a, WgState<BLOCK>& S, uint32_t L, uint64_t ch, uint32_t& mn,
                              uint32_t& sn, bool local) {
    const uint32_t jb = (uint32_t)ch, je = (uint32_t)(ch >> 32);
    const uint32_t lane = lane_id();
    // unconditional adjacency loads, as in process_vertices (inactive lanes read arc jb < je)
    bool act[kU];
    uint32_t u[kU];
#pragma unroll
    for (uint32_t k = 0; k < kU; ++k) {
        act[k] = jb + k * 64 + lane < je;
        u[k] = a.nbr[act[k] ? jb + k * 64 + lane : jb];
    }
    for (uint32_t base = jb; base < je; base += 64 * kU) {
        const bool more = base + 64 * kU < je;
        bool actn[kU];
        uint32_t un[kU];
#pragma unroll
        for (uint32_t k = 0; k < kU; ++k) {
            const uint32_t jn = base + 64 * kU + k * 64 + lane;
            actn[k] = jn < je;
            un[k] = 0;
        }
        if (PREFETCH && more) {
#pragma unroll
            for (uint32_t k = 0; k < kU; ++k) un[k] = a.nbr[actn[k] ? base + 64 * kU + k * 64 + lane : jb];
        }
        relax_group(a, S, L, act, u, mn, sn, local);
        if (!PREFETCH && more) {
#pragma unroll
            for (uint32_t k = 0; k < kU; ++k) un[k] = a.nbr[actn[k] ? base + 64 * kU + k * 64 + lane : jb];
        }
#pragma unroll
        for (uint32_t k = 0; k < kU; ++k) {
            u[k] = un[k];
            act[k] = actn[k];
        }
    }
}

// ---------------------------------------------------------------------------------------
// select: stage every v with core[v] == L (the working copy of the degrees is made by a
// separate kernel before the peel); the flushes record them in order[] and chunk heavy ones.
template <int BLOCK>
__device__ __forceinline__ void select_quad(const PeelArgs& a, WgState<BLOCK>& S, uint32_t L, uint4 d,
                                            uint32_t v0, uint32_t& mn, uint32_t& sn) {
    const uint32_t m0 = d.x == L, m1 = d.y == L, m2 = d.z == L, m3 = d.w == L;
    if (d.x > L && d.x < mn) mn = d.x;
    if (d.y > L && d.y < mn) mn = d.y;
    if (d.z > L && d.z < mn) mn = d.z;
    if (d.w > L && d.w < mn) mn = d.w;
    const uint32_t c = m0 + m1 + m2 + m3;
    const uint64_t b0 = __ballot(c & 1), b1 = __ballot(c & 2), b2 = __ballot(c & 4);
    if ((b0 | b1 | b2) == 0) return;
    const uint32_t total = __popcll(b0) + 2 * __popcll(b1) + 4 * __popcll(b2);
    if (sn + total > kStageCap) flush_stage(a, S, sn, false, false);
    uint32_t p = sn + rank_below(b0) + 2 * rank_below(b1) + 4 * rank_below(b2);
    uint32_t* st = S.stage[threadIdx.x / 64];
    if (m0) st[p++] = v0;
    if (m1) st[p++] = v0 + 1;
    if (m2) st[p++] = v0 + 2;
    if (m3) st[p++] = v0 + 3;
    sn += total;
    __builtin_amdgcn_wave_barrier();
}

template <int BLOCK>
__device__ void select_level(const PeelArgs& a, WgState<BLOCK>& S, uint32_t L, uint32_t& mn) {
    constexpr uint32_t WPB = BLOCK / 64;
    constexpr uint32_t U = 4;  // 16-byte loads in flight per lane
    const uint32_t lane = lane_id();
    const uint32_t gw = S.vbid * WPB + threadIdx.x / 64;
    const uint32_t GW = S.npart * WPB;
    const __amdgpu_buffer_rsrc_t core = make_rsrc(a.core, a.nvec * 16u);
    uint32_t sn = 0;
    // wave gw owns the 64*U-quad tiles gw, gw + GW, ...; lanes read consecutive quads
    for (uint32_t t = gw * 64 * U; t < a.nvec; t += GW * 64 * U) {
        uint4 d[U];
#pragma unroll
        for (uint32_t k = 0; k < U; ++k) {
            const uint32_t i = t + k * 64 + lane;
            d[k] = i < a.nvec ? load_quad_coherent(core, i) : make_uint4(kSent, kSent, kSent, kSent);
        }
#pragma unroll
        for (uint32_t k = 0; k < U; ++k) select_quad(a, S, L, d[k], 4 * (t + k * 64 + lane), mn, sn);
    }
    flush_stage(a, S, sn, false, false);
}

// One sub-round for this workgroup: its static share of chunks and vertex batches, then up to
// lq_rounds local rounds over the crossings it kept in LDS.
template <int BLOCK, bool PREFETCH>
__device__ void process_subround(const PeelArgs& a, WgState<BLOCK>& S, uint32_t L, uint32_t src,
                                 uint32_t vbeg, uint32_t vend, uint32_t cbeg, uint32_t cend,
                                 uint32_t& mn) {
    constexpr uint32_t WPB = BLOCK / 64;
    const uint32_t lane = lane_id();
    const uint32_t gw = S.vbid * WPB + threadIdx.x / 64;
    const uint32_t GW = S.npart * WPB;
    const uint32_t* vsrc = src ? a.work : a.order;
    if (threadIdx.x == 0) {
        S.cur_n = 0;
        S.head = 0;
        S.nxt_n = 0;
        S.larcs = 0;
        S.cur = 0;
    }
    __syncthreads();
    bool local = a.lq_depth != 0 && a.lq_rounds != 0;
    uint32_t sn = 0;
    // chunks are equal-sized: dealt round-robin over every wave of the grid
    for (uint32_t c = cbeg + gw; c < cend; c += GW) {
        uint64_t ch = 0;
        if (lane == 0) ch = ld_rlx64(a.chunks + c);
        ch = (uint64_t)bcast_first((uint32_t)ch) | ((uint64_t)bcast_first((uint32_t)(ch >> 32)) << 32);
        process_chunk<BLOCK, PREFETCH>(a, S, L, ch, mn, sn, local);
    }
    for (uint32_t b = vbeg + 64 * gw; b < vend; b += 64 * GW) {
        const bool mine = b + lane < vend;
        const uint32_t v = mine ? ld_rlx(vsrc + b + lane) : 0;
        process_vertices<BLOCK, PREFETCH>(a, S, L, mine, v, mn, sn, local);
    }
    flush_stage(a, S, sn, true, local);
    for (uint32_t r = 0;; ++r) {
        __syncthreads();
        const uint32_t nn = S.nxt_n;
        __syncthreads();
        if (nn == 0) break;
        if (threadIdx.x == 0) {
            S.cur ^= 1;
            S.cur_n = nn < a.lq_depth ? nn : a.lq_depth;
            S.head = 0;
            S.nxt_n = 0;
            S.larcs = 0;
        }
        __syncthreads();
        local = r + 1 < a.lq_rounds;  // the last local round spills whatever it uncovers
        const uint32_t cn = S.cur_n;
        while (true) {
            uint32_t i = 0;
            if (lane == 0) i = __hip_atomic_fetch_add(&S.head, 64u, __ATOMIC_RELAXED, KC_WG);
            i = bcast_first(i);
            if (i >= cn) break;
            const bool mine = i + lane < cn;
            const uint32_t v = mine ? S.lq[S.cur][i + lane] : 0;
            process_vertices<BLOCK, PREFETCH>(a, S, L, mine, v, mn, sn, local);
        }
        flush_stage(a, S, sn, true, local);
    }
}

// ---------------------------------------------------------------------------------------
// Barrier hooks: run by every thread of the last workgroup to arrive; only thread 0 writes.

template <int BLOCK>
__device__ __forceinline__ void level_end(const PeelArgs& a, WgState<BLOCK>& S, uint32_t L, uint32_t* snap) {
    // smallest degree above L left anywhere: reduce the per-workgroup minima
    uint32_t m = kSent;
    for (uint32_t i = threadIdx.x; i < S.npart; i += BLOCK) m = min(m, ld_rlx(a.wgmin + i));
    m = wave_min(m);
    if (lane_id() == 0) S.red[threadIdx.x / 64] = m;
    __syncthreads();
    if (threadIdx.x != 0) return;
    for (int w = 0; w < BLOCK / 64; ++w) m = min(m, S.red[w]);
    const uint32_t t = ld_rlx(cw(a, C_OTAIL));
    uint32_t phase = PH_NEXT_LEVEL, next = L + 1;
    if (!(a.flags & FLAG_NO_LEVEL_SKIP) && m != kSent && m > next) next = m;
    add_rlx(cw(a, C_LEVELS), 1u);
    if (t >= a.V) {
        a.shell[L + 1] = t;
        phase = PH_DONE;
        next = L;
    } else if (next > a.dmax) {
        __hip_atomic_fetch_max(cw(a, C_ERR), 2u, __ATOMIC_RELAXED, KC_AGENT);
        phase = PH_DONE;
    } else {
        for (uint32_t k = L + 1; k <= next; ++k) a.shell[k] = t;
        st_rlx(cw(a, C_LSTART), t);
    }
    st_rlx(snap + S_PHASE, phase);
    st_rlx(snap + S_NEXT, next);
    trace_rec(a, TR_LEVEL, 0, L);
}

// after the select of level L
template <int BLOCK>
__device__ __forceinline__ void hook_after_scan(const PeelArgs& a, WgState<BLOCK>& S, uint32_t L, uint32_t* snap) {
    uint32_t vbeg = 0, vend = 0, cbeg = 0, cend = 0;
    if (threadIdx.x == 0) {
        vbeg = ld_rlx(cw(a, C_LSTART));
        vend = ld_rlx(cw(a, C_OTAIL));
        cbeg = ld_rlx(cw(a, C_CSTART));
        cend = ld_rlx(cw(a, C_CTAIL));
        S.red[0] = vbeg == vend && cbeg == cend;
    }
    __syncthreads();
    if (S.red[0]) {  // empty frontier (only without level skipping)
        __syncthreads();
        level_end(a, S, L, snap);
        return;
    }
    if (threadIdx.x != 0) return;
    st_rlx(cw(a, C_WSTART), ld_rlx(cw(a, C_WTAIL)));
    st_rlx(cw(a, C_CSTART), cend);
    st_rlx(snap + S_PHASE, PH_SUBROUND);
    st_rlx(snap + S_SRC, 0);
    st_rlx(snap + S_VBEG, vbeg);
    st_rlx(snap + S_VEND, vend);
    st_rlx(snap + S_CBEG, cbeg);
    st_rlx(snap + S_CEND, cend);
    trace_rec(a, TR_SCAN, 0, L);
}

template <int BLOCK>
__device__ __forceinline__ void hook_after_subround(const PeelArgs& a, WgState<BLOCK>& S, uint32_t L, uint32_t* snap) {
    uint32_t wbeg = 0, wend = 0, cbeg = 0, cend = 0;
    if (threadIdx.x == 0) {
        wbeg = ld_rlx(cw(a, C_WSTART));
        wend = ld_rlx(cw(a, C_WTAIL));
        cbeg = ld_rlx(cw(a, C_CSTART));
        cend = ld_rlx(cw(a, C_CTAIL));
        S.red[0] = wbeg == wend && cbeg == cend;
        const uint32_t n = add_rlx(cw(a, C_SUBROUNDS), 1u);
        trace_rec(a, TR_SUB, n, L);
    }
    __syncthreads();
    if (S.red[0]) {
        __syncthreads();
        level_end(a, S, L, snap);
        return;
    }
    if (threadIdx.x != 0) return;
    st_rlx(cw(a, C_WSTART), wend);
    st_rlx(cw(a, C_CSTART), cend);
    st_rlx(snap + S_PHASE, PH_SUBROUND);
    st_rlx(snap + S_SRC, 1);
    st_rlx(snap + S_VBEG, wbeg);
    st_rlx(snap + S_VEND, wend);
    st_rlx(snap + S_CBEG, cbeg);
    st_rlx(snap + S_CEND, cend);
}

template <int BLOCK, bool PREFETCH>
__global__ __launch_bounds__(BLOCK) void peel_persistent(PeelArgs a) {
    __shared__ WgState<BLOCK> S;
    // Every call below is inlined at its call site. An out-of-line callee taking the arguments
    // by reference would keep them in scratch memory, reloaded from there at every use.
    bool registered;
    [[clang::always_inline]] registered = register_workgroup<BLOCK>(a, S);
    if (!registered) return;
    uint32_t L = 0, nb = 0;
    bool scan = true;
    uint32_t src = 0, vbeg = 0, vend = 0, cbeg = 0, cend = 0;
    uint32_t mn = kSent;
    if (threadIdx.x == 0) S.mn = kSent;
    __syncthreads();
    while (true) {
        uint32_t* snap = cw(a, C_SNAP + (nb & 1));
        if (threadIdx.x == 0) {
            st_rlx(a.dbg + S.vbid * 8 + 0, nb);
            st_rlx(a.dbg + S.vbid * 8 + 1, L);
            st_rlx(a.dbg + S.vbid * 8 + 2, scan ? 1u : 2u);
        }
        if (scan) {
            [[clang::always_inline]] select_level<BLOCK>(a, S, L, mn);
        } else {
            [[clang::always_inline]] process_subround<BLOCK, PREFETCH>(a, S, L, src, vbeg, vend, cbeg, cend, mn);
        }
        // publish this workgroup's smallest degree above L seen so far in the level
        mn = wave_min(mn);
        if (lane_id() == 0) __hip_atomic_fetch_min(&S.mn, mn, __ATOMIC_RELAXED, KC_WG);
        __syncthreads();
        if (threadIdx.x == 0) st_rlx(a.wgmin + S.vbid, S.mn);
        if (scan)
            [[clang::always_inline]] grid_barrier<BLOCK>(a, S, [&] { hook_after_scan<BLOCK>(a, S, L, snap); });
        else
            [[clang::always_inline]] grid_barrier<BLOCK>(a, S, [&] { hook_after_subround<BLOCK>(a, S, L, snap); });
        nb++;
        if (ld_rlx(cw(a, C_ERR))) break;
        const uint32_t phase = ld_rlx(snap + S_PHASE);
        if (threadIdx.x == 0) st_rlx(a.dbg + S.vbid * 8 + 3, phase);
        if (phase == PH_SUBROUND) {
            scan = false;
            src = ld_rlx(snap + S_SRC);
            vbeg = ld_rlx(snap + S_VBEG);
            vend = ld_rlx(snap + S_VEND);
            cbeg = ld_rlx(snap + S_CBEG);
            cend = ld_rlx(snap + S_CEND);
        } else if (phase == PH_NEXT_LEVEL) {
            L = ld_rlx(snap + S_NEXT);
            scan = true;
            mn = kSent;
            __syncthreads();
            if (threadIdx.x == 0) S.mn = kSent;
            __syncthreads();
        } else {
            break;
        }
    }
    if (S.vbid == 0 && threadIdx.x == 0) {
        uint32_t* r = cw(a, C_RESULT);
        r[R_ERR] = ld_rlx(cw(a, C_ERR));
        r[R_KMAX] = L;
        r[R_LEVELS] = ld_rlx(cw(a, C_LEVELS));
        r[R_SUBROUNDS] = ld_rlx(cw(a, C_SUBROUNDS));
        r[R_NPART] = S.npart;
        r[R_NTRACE] = ld_rlx(cw(a, C_NTRACE));
        r[R_WORDS + 3] = ld_rlx(cw(a, C_CTAIL));
    }
}

}  // namespace kcdev
