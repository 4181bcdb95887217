// stepwise_kernel.h — the kernels of the reference-shaped variant (KC_VARIANT_STEPWISE).
#pragma once
#include "device_common.h"

namespace kcdev {

// ---- the reference-shaped variant (KC_VARIANT_STEPWISE): the host runs the level loop
// (src/ours.cc:117-160), launching a select kernel and then process kernels per level and
// reading the frontier size back after each, as the reference reads its count per level.

// select (src/ours.cc:2-33): append every v with core[v] == L to order[] at ctr[0]; ctr[1] gets
// the smallest degree above L. Plain loads: nothing changes core[] during this launch.
__global__ void step_select(const uint32_t* core, uint32_t V, uint32_t L, uint32_t* order, uint32_t* ctr) {
    uint32_t mn = kSent;
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t b = blockIdx.x * blockDim.x; b < V; b += stride) {
        const uint32_t v = b + threadIdx.x;
        const uint32_t d = v < V ? core[v] : kSent;
        if (d > L && d < mn) mn = d;
        const uint64_t m = __ballot(v < V && d == L);
        if (m) {
            uint32_t base = 0;
            if (lane_id() == 0) base = add_rlx(ctr, (uint32_t)__popcll(m));
            base = bcast_first(base);
            if (v < V && d == L) order[base + rank_below(m)] = v;
        }
    }
    mn = wave_min(mn);
    if (lane_id() == 0 && mn != kSent) __hip_atomic_fetch_min(ctr + 1, mn, __ATOMIC_RELAXED, KC_AGENT);
}

// process (src/ours.cc:38-114): one wave per frontier vertex of order[beg, end); its lanes walk
// the adjacency; u with core[u] > L is decremented, a == L+1 appends u at ctr[0], a <= L undoes.
__global__ void step_process(const uint32_t* off, const uint32_t* nbr, uint32_t* core, uint32_t* order,
                            uint32_t beg, uint32_t end, uint32_t L, uint32_t* ctr) {
    const uint32_t lane = lane_id();
    const uint32_t gw = (blockIdx.x * blockDim.x + threadIdx.x) / 64;
    const uint32_t GW = gridDim.x * blockDim.x / 64;
    uint32_t mn = kSent;
    for (uint32_t i = beg + gw; i < end; i += GW) {
        const uint32_t v = order[i];
        const uint32_t s = off[v], e = off[v + 1];
        for (uint32_t j0 = s; j0 < e; j0 += 64) {
            const uint32_t j = j0 + lane;
            bool cross = false;
            uint32_t u = 0;
            if (j < e) {
                u = nbr[j];
                if (core[u] > L) {
                    const uint32_t a = __hip_atomic_fetch_sub(core + u, 1u, __ATOMIC_RELAXED, KC_AGENT);
                    if (a == L + 1) {
                        cross = true;
                    } else if (a <= L) {
                        __hip_atomic_fetch_add(core + u, 1u, __ATOMIC_RELAXED, KC_AGENT);
                    } else if (a - 1 < mn) {
                        mn = a - 1;
                    }
                }
            }
            const uint64_t m = __ballot(cross);
            if (m) {
                uint32_t base = 0;
                if (lane == 0) base = add_rlx(ctr, (uint32_t)__popcll(m));
                base = bcast_first(base);
                if (cross) order[base + rank_below(m)] = u;
            }
        }
    }
    mn = wave_min(mn);
    if (lane == 0 && mn != kSent) __hip_atomic_fetch_min(ctr + 1, mn, __ATOMIC_RELAXED, KC_AGENT);
}

}  // namespace kcdev
