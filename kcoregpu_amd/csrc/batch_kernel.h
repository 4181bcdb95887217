// batch_kernel.h — batched peeling: many independent graphs packed into one CSR, peeled by one
// persistent kernel in which every workgroup owns one graph at a time.
//
// Same method as the single-graph kernel (reference src/ours.cc:2-160, per graph), but graphs are
// independent, so a workgroup runs a graph's whole level loop with workgroup barriers only: no
// grid barrier, no cross-workgroup queue. Workgroups claim graphs from a device counter in the
// order the host gives (largest first, so the long graphs start early), each with its own level
// counter.
//
// Within a graph, each round's frontier (a range of the graph's order[] segment) is processed in
// tiles of BLOCK vertices: a workgroup-wide prefix sum of their degrees spreads the tile's arcs
// evenly over every thread of the workgroup (a hub's arcs over all waves: the reference's
// block-per-vertex fallback, applied to whole tiles), with kU arcs per thread in flight.
// Crossings are appended to the graph's order[] through an LDS counter and form the next round.
// The next level is the smallest degree left above the level, so empty levels are not scanned.
//
// Packed layout (include/kcore.h kc_batch_csr): graph g owns vertices [vtx_off[g], vtx_off[g+1])
// of the packed vertex arrays (local ids 0..V_g-1), arcs [arc_off[g], arc_off[g+1]) of the packed
// neighbour array (local ids), and V_g+1 local offsets starting at off[vtx_off[g] + g].
#pragma once
#include "device_common.h"

namespace kcdev {

struct BatchArgs {
    const uint64_t* vtx_off;  // G+1
    const uint64_t* arc_off;  // G+1
    const uint32_t* off;      // sum(V_g + 1): local offsets
    const uint32_t* nbr;      // local neighbour ids
    uint32_t* core;           // packed working degrees -> core numbers (copied from the input
                              // degrees by a separate kernel before the launch)
    uint32_t* order;          // packed per-graph peel order
    uint32_t* kmax;           // G
    const uint32_t* sched;    // G graph ids, in claim order
    uint32_t* ctrl;           // [0] next claim, [64] levels, [128] rounds, [192] error,
                              // [224..225] degree entries scanned (uint64)
    uint32_t G;
    uint32_t flags;
};

constexpr uint32_t kBU = 4;  // arcs per thread in flight

template <int BLOCK>
struct BatchState {
    static constexpr int WPB = BLOCK / 64;
    uint32_t tincl[BLOCK];  // tile: inclusive degree prefix
    uint32_t tstart[BLOCK]; // tile: adjacency start
    uint32_t wtot[WPB + 1];
    uint32_t tail;          // order[] entries of this graph so far
    uint32_t mn;            // smallest degree above the level
    uint32_t g;             // current graph
};

// Workgroup-wide inclusive prefix sum (all threads call); *total receives the sum.
template <int BLOCK>
__device__ __forceinline__ uint32_t block_incl_scan(BatchState<BLOCK>& S, uint32_t x, uint32_t* total) {
    constexpr uint32_t WPB = BLOCK / 64;
    const uint32_t lane = lane_id(), w = threadIdx.x / 64;
    const uint32_t inc = wave_incl_scan(x);
    if (lane == 63) S.wtot[w] = inc;
    __syncthreads();
    if (w == 0) {
        const uint32_t t = lane < WPB ? S.wtot[lane] : 0;
        const uint32_t ti = wave_incl_scan(t);
        if (lane < WPB) S.wtot[lane] = ti - t;
        if (lane == WPB - 1) S.wtot[WPB] = ti;
    }
    __syncthreads();
    *total = S.wtot[WPB];
    return inc + S.wtot[w];
}

// Append the lanes' u (where pred) to the graph's order[] (LDS tail, one atomic per wave).
template <int BLOCK>
__device__ __forceinline__ void batch_append(BatchState<BLOCK>& S, uint32_t* order_g, bool pred, uint32_t u) {
    const uint64_t m = __ballot(pred);
    if (m == 0) return;
    uint32_t base = 0;
    if (lane_id() == 0) base = __hip_atomic_fetch_add(&S.tail, (uint32_t)__popcll(m), __ATOMIC_RELAXED, KC_WG);
    base = bcast_first(base);
    if (pred) st_rlx(order_g + base + rank_below(m), u);
}

// Peel one graph with this workgroup: off[V+1], nbr (local ids), core[V] (already holding the
// input degrees, with no stale copy in any cache: see the callers), order[V] (peel order).
// Returns kmax; *levels / *rounds count scanned levels and processing rounds.
template <int BLOCK>
__device__ uint32_t wg_peel_graph(BatchState<BLOCK>& S, const uint32_t* off, const uint32_t* nbr,
                                  uint32_t* core, uint32_t* order, uint32_t Vg, uint32_t flags,
                                  uint32_t* levels_out, uint32_t* rounds_out, uint32_t* err) {
    const uint32_t tid = threadIdx.x;
    if (tid == 0) {
        S.tail = 0;
        S.mn = kSent;
    }
    __syncthreads();
    uint32_t L = 0, levels = 0, rounds = 0;
    // select reads the degrees as 16-byte device-coherent (SC1) buffer loads, which see the
    // other waves' atomics; the graph's array may start inside a 16-byte quad (shift words)
    const uintptr_t cb = reinterpret_cast<uintptr_t>(core);
    const uint32_t shift = (uint32_t)((cb & 15u) >> 2);
    const uint64_t nq64 = ((uint64_t)Vg + shift + 3) / 4;
    const bool quads = nq64 * 16 < (1ull << 31);
    const uint32_t nq = quads ? (uint32_t)nq64 : 0;
    const __amdgpu_buffer_rsrc_t rs = make_rsrc(reinterpret_cast<const void*>(cb & ~(uintptr_t)15), nq * 16u);
    while (true) {
        // select: every v with core[v] == L
        uint32_t mn = kSent;
        const uint32_t lstart = S.tail;
        if (quads) {
            for (uint32_t q0 = 0; q0 < nq; q0 += BLOCK * kBU) {
                uint4 d[kBU];
#pragma unroll
                for (uint32_t k = 0; k < kBU; ++k) {
                    const uint32_t q = q0 + k * BLOCK + tid;
                    d[k] = q < nq ? load_quad_coherent(rs, q) : make_uint4(kSent, kSent, kSent, kSent);
                }
#pragma unroll
                for (uint32_t k = 0; k < kBU; ++k) {
                    // vertex of the quad's first word; words before the graph wrap to >= Vg
                    const uint32_t v0 = 4u * (q0 + k * BLOCK + tid) - shift;
                    const uint32_t x[4] = {d[k].x, d[k].y, d[k].z, d[k].w};
                    uint32_t c = 0;
#pragma unroll
                    for (uint32_t j = 0; j < 4; ++j) {
                        const bool in = v0 + j < Vg;
                        if (in && x[j] > L && x[j] < mn) mn = x[j];
                        c += (in && x[j] == L) ? 1u : 0u;
                    }
                    // one reservation per wave for all of its matches in this quad row
                    if (__ballot(c != 0)) {
                        const uint32_t inc = wave_incl_scan(c);
                        uint32_t base = 0;
                        if (lane_id() == 63) base = __hip_atomic_fetch_add(&S.tail, inc, __ATOMIC_RELAXED, KC_WG);
                        base = (uint32_t)__builtin_amdgcn_readlane((int)base, 63);
                        uint32_t pos = base + inc - c;
#pragma unroll
                        for (uint32_t j = 0; j < 4; ++j)
                            if (v0 + j < Vg && x[j] == L) st_rlx(order + pos++, v0 + j);
                    }
                }
            }
        } else {
            for (uint32_t i0 = 0; i0 < Vg; i0 += BLOCK * kBU) {
                uint32_t d[kBU];
#pragma unroll
                for (uint32_t k = 0; k < kBU; ++k) {
                    const uint32_t i = i0 + k * BLOCK + tid;
                    d[k] = i < Vg ? ld_rlx(core + i) : kSent;
                }
#pragma unroll
                for (uint32_t k = 0; k < kBU; ++k) {
                    if (d[k] > L && d[k] < mn) mn = d[k];
                    batch_append(S, order, d[k] == L, i0 + k * BLOCK + tid);
                }
            }
        }
        levels++;
        drain_memory();
        __syncthreads();
        // rounds: process [h, t), crossings append at t and form the next round
        uint32_t h = lstart;
        uint32_t t = S.tail;
        while (h < t) {
            rounds++;
            for (uint32_t tb = h; tb < t; tb += BLOCK) {
                const uint32_t i = tb + tid;
                uint32_t s = 0, dv = 0;
                if (i < t) {
                    const uint32_t v = ld_rlx(order + i);
                    s = off[v];
                    dv = off[v + 1] - s;
                }
                uint32_t T;
                const uint32_t inc = block_incl_scan(S, dv, &T);
                S.tincl[tid] = inc;
                S.tstart[tid] = s;
                __syncthreads();
                for (uint32_t jb = 0; jb < T; jb += BLOCK * kBU) {
                    // kBU searches advanced step by step together (independent LDS reads per
                    // step), then unconditional adjacency loads: inactive lanes take arc 0,
                    // which exists since T > 0
                    bool act[kBU];
                    uint32_t u[kBU], jj[kBU], o[kBU];
#pragma unroll
                    for (uint32_t k = 0; k < kBU; ++k) {
                        const uint32_t j = jb + k * BLOCK + tid;
                        act[k] = j < T;
                        jj[k] = act[k] ? j : 0;
                        o[k] = 0;
                    }
#pragma unroll
                    for (uint32_t step = BLOCK / 2; step; step >>= 1) {
#pragma unroll
                        for (uint32_t k = 0; k < kBU; ++k)
                            if (S.tincl[o[k] + step - 1] <= jj[k]) o[k] += step;
                    }
#pragma unroll
                    for (uint32_t k = 0; k < kBU; ++k) {
                        const uint32_t ex = o[k] ? S.tincl[o[k] - 1] : 0;
                        u[k] = nbr[S.tstart[o[k]] + (jj[k] - ex)];
                    }
                    // prechecks loaded unconditionally so they overlap (inactive lanes read
                    // the graph's first offset, read-only), as in the single-graph kernels
                    uint32_t du[kBU], old[kBU];
#pragma unroll
                    for (uint32_t k = 0; k < kBU; ++k) {
                        const uint32_t d = *(act[k] ? core + u[k] : off);
                        du[k] = act[k] ? d : 0;
                    }
#pragma unroll
                    for (uint32_t k = 0; k < kBU; ++k)
                        old[k] = du[k] > L ? __hip_atomic_fetch_sub(core + u[k], 1u, __ATOMIC_RELAXED, KC_AGENT) : 0;
#pragma unroll
                    for (uint32_t k = 0; k < kBU; ++k) {
                        const bool cross = du[k] > L && old[k] == L + 1;
                        if (du[k] > L && old[k] <= L)
                            __hip_atomic_fetch_add(core + u[k], 1u, __ATOMIC_RELAXED, KC_AGENT);
                        else if (du[k] > L && old[k] > L + 1 && old[k] - 1 < mn)
                            mn = old[k] - 1;
                        batch_append(S, order, cross, u[k]);
                    }
                }
                __syncthreads();  // the tile arrays are rewritten by the next tile
            }
            drain_memory();
            __syncthreads();
            h = t;
            t = S.tail;
        }
        // level end
        mn = wave_min(mn);
        if (lane_id() == 0) __hip_atomic_fetch_min(&S.mn, mn, __ATOMIC_RELAXED, KC_WG);
        __syncthreads();
        const uint32_t tl = S.tail;
        const uint32_t m = S.mn;
        __syncthreads();
        if (tl >= Vg) break;
        uint32_t next = L + 1;
        if (!(flags & 1u) && m != kSent && m > next) next = m;
        if (next >= Vg) {  // a core number is below V_g on a simple graph: no progress
            if (tid == 0) *err = 2;
            break;
        }
        L = next;
        if (tid == 0) S.mn = kSent;
        __syncthreads();
    }
    *levels_out = levels;
    *rounds_out = rounds;
    return Vg ? L : 0;
}

template <int BLOCK>
__device__ void batch_peel_graph(const BatchArgs& a, BatchState<BLOCK>& S, uint32_t g) {
    const uint64_t vb = a.vtx_off[g];
    const uint32_t Vg = (uint32_t)(a.vtx_off[g + 1] - vb);
    // the working degrees were copied from the input degrees by a kernel launched before this
    // one, so within this launch a graph's degrees only ever decrease (plain prechecks may read
    // stale values, but never too-low ones)
    uint32_t levels = 0, rounds = 0, err = 0;
    const uint32_t km = wg_peel_graph<BLOCK>(S, a.off + vb + g, a.nbr + a.arc_off[g], a.core + vb, a.order + vb,
                                             Vg, a.flags, &levels, &rounds, &err);
    if (threadIdx.x == 0) {
        st_rlx(a.kmax + g, km);
        add_rlx(a.ctrl + 64, levels);
        add_rlx(a.ctrl + 128, rounds);
        // degree entries scanned: V_g per level this graph ran (words 224-225, one uint64)
        __hip_atomic_fetch_add(reinterpret_cast<uint64_t*>(a.ctrl + 224), (uint64_t)levels * Vg, __ATOMIC_RELAXED,
                               KC_AGENT);
        if (err) __hip_atomic_fetch_max(a.ctrl + 192, err, __ATOMIC_RELAXED, KC_AGENT);
    }
}

template <int BLOCK>
__global__ __launch_bounds__(BLOCK) void peel_batch(BatchArgs a) {
    __shared__ BatchState<BLOCK> S;
    while (true) {
        if (threadIdx.x == 0) {
            const uint32_t c = add_rlx(a.ctrl, 1u);
            S.g = c < a.G ? a.sched[c] : kSent;
        }
        __syncthreads();
        const uint32_t g = S.g;
        __syncthreads();
        if (g == kSent) break;
        batch_peel_graph<BLOCK>(a, S, g);
        __syncthreads();
    }
}

}  // namespace kcdev
