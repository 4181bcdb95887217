// device_common.h — wave64 / CDNA4 primitives shared by the peeling kernels.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#define KC_AGENT __HIP_MEMORY_SCOPE_AGENT
#define KC_WG __HIP_MEMORY_SCOPE_WORKGROUP

namespace kcdev {

constexpr uint32_t kSent = 0xFFFFFFFFu;
constexpr uint64_t kSent64 = 0xFFFFFFFFFFFFFFFFull;

__device__ __forceinline__ uint32_t lane_id() {
    return __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u));
}
// number of set bits of m in the lanes below this one
__device__ __forceinline__ uint32_t rank_below(uint64_t m) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}
__device__ __forceinline__ uint32_t bcast_first(uint32_t x) { return __builtin_amdgcn_readfirstlane(x); }
__device__ __forceinline__ uint32_t read_lane(uint32_t x, int lane) {
    return (uint32_t)__builtin_amdgcn_readlane((int)x, lane);
}

__device__ __forceinline__ uint32_t ld_rlx(const uint32_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, KC_AGENT);
}
__device__ __forceinline__ uint64_t ld_rlx64(const uint64_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, KC_AGENT);
}
__device__ __forceinline__ void st_rlx(uint32_t* p, uint32_t v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, KC_AGENT);
}
__device__ __forceinline__ uint32_t add_rlx(uint32_t* p, uint32_t v) {
    return __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, KC_AGENT);
}
__device__ __forceinline__ void st_rlx64(uint64_t* p, uint64_t v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, KC_AGENT);
}

// Device-coherent 16-byte loads: a buffer load with the SC1 cache-policy bit reads at device
// scope, so it sees values that workgroups on other XCDs changed with atomics, without an
// L2 invalidation. Loads past num_records return zeros.
using u32x4 = __attribute__((ext_vector_type(4))) unsigned int;
constexpr int kCpolSc1 = 16;
__device__ __forceinline__ __amdgpu_buffer_rsrc_t make_rsrc(const void* base, uint32_t bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), 0, (int)bytes, 0x00020000);
}
__device__ __forceinline__ uint4 load_quad_coherent(__amdgpu_buffer_rsrc_t r, uint32_t quad) {
    const u32x4 v = __builtin_amdgcn_raw_buffer_load_b128(r, (int)(quad * 16u), 0, kCpolSc1);
    return make_uint4(v.x, v.y, v.z, v.w);
}

// Wait until every memory operation this wave issued has completed. Data that other
// workgroups read is written with device-scope (SC1) stores, so completion is visibility:
// this replaces an agent-scope release fence, whose L2 write-back costs tens of microseconds.
__device__ __forceinline__ void drain_memory() { __builtin_amdgcn_s_waitcnt(0); }

// Wave64 inclusive prefix sum with DPP: four row_shr steps scan each row of 16 lanes, then
// row_bcast:15 and row_bcast:31 carry row totals into the following rows.
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

// Minimum over the wave (all lanes active); every lane receives it.
__device__ __forceinline__ uint32_t wave_min(uint32_t x) {
    for (int o = 32; o; o >>= 1) x = min(x, (uint32_t)__shfl_xor((int)x, o));
    return x;
}

// Reserve `cnt` consecutive slots of a global append counter for the wave (all lanes call
// with the same cnt); returns the first slot in every lane.
__device__ __forceinline__ uint32_t wave_reserve(uint32_t* counter, uint32_t cnt) {
    uint32_t base = 0;
    if (lane_id() == 0) base = add_rlx(counter, cnt);
    return bcast_first(base);
}

// Claim up to k items of a range ending at `end` from a claim counter. Returns `end` when the
// range is exhausted; the counter is only incremented while it is below `end`.
__device__ __forceinline__ uint32_t wave_claim(uint32_t* counter, uint32_t k, uint32_t end) {
    uint32_t r = end;
    if (lane_id() == 0) {
        if (ld_rlx(counter) < end) r = add_rlx(counter, k);
        if (r > end) r = end;
    }
    return bcast_first(r);
}

// Time-limited wait bookkeeping: every spin loop calls expired(); once the wall-clock budget
// of one wait is spent the error flag is raised, and every wait of every wave then ends, so
// the grid always drains.
struct Waiter {
    uint64_t t0 = 0;
    uint32_t n = 0;
    __device__ __forceinline__ bool expired(uint32_t* err, uint64_t limit) {
        if ((++n & 31) != 0) return false;
        uint64_t now = wall_clock64();
        if (t0 == 0) t0 = now;
        if (now - t0 > limit) {
            __hip_atomic_fetch_max(err, 1u, __ATOMIC_RELAXED, KC_AGENT);
            return true;
        }
        return ld_rlx(err) != 0;
    }
};

}  // namespace kcdev
