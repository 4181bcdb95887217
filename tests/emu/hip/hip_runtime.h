// CPU stand-in for <hip/hip_runtime.h>, used only by the wave64 emulator in tests/emu/ to run
// this project's device-code headers on the host. Each workgroup runs on one OS thread with one
// fiber per work-item; wave-wide operations (ballot, DPP, readlane, shuffles, wave barrier) and
// workgroup barriers are synchronisation points at which the emulator (emu_sched.h) lines the
// lanes up. LDS (__shared__) is thread_local, i.e. per workgroup. Memory is host memory; atomics
// map to the compiler's __atomic builtins. This is test infrastructure, not a HIP runtime.
//
// Weak-memory mode (environment KCEMU_WEAK=1): a work-item's relaxed atomic stores wait in its own
// store buffer, invisible to every other work-item, until it drains them (s_waitcnt, a fence, a
// release store, an atomic read-modify-write of a buffered address) or the scheduler lets some of
// them complete at a random later point; its own loads see them. Stores to different addresses
// thus become visible in any order and late, as in-flight stores on the GPU may, and a kernel that
// publishes data without draining first is caught. Workgroup barriers drain nothing.
#pragma once
#include <cstddef>
#include <cstdint>
#include <cstring>

#define __global__
#define __device__
#define __host__
#define __forceinline__ inline __attribute__((always_inline))
#define __launch_bounds__(...)
#define __shared__ static thread_local

#define __HIP_MEMORY_SCOPE_SINGLETHREAD 1
#define __HIP_MEMORY_SCOPE_WAVEFRONT 2
#define __HIP_MEMORY_SCOPE_WORKGROUP 3
#define __HIP_MEMORY_SCOPE_AGENT 4
#define __HIP_MEMORY_SCOPE_SYSTEM 5

struct uint4 {
    uint32_t x, y, z, w;
};
inline uint4 make_uint4(uint32_t x, uint32_t y, uint32_t z, uint32_t w) { return uint4{x, y, z, w}; }
struct kcemu_dim3 {
    uint32_t x, y, z;
};

inline uint32_t min(uint32_t a, uint32_t b) { return a < b ? a : b; }
inline uint32_t max(uint32_t a, uint32_t b) { return a > b ? a : b; }
inline uint64_t min(uint64_t a, uint64_t b) { return a < b ? a : b; }
inline int min(int a, int b) { return a < b ? a : b; }

namespace kcemu {
// provided by emu_sched.h
kcemu_dim3 tid3();
kcemu_dim3 bid3();
kcemu_dim3 bdim3();
kcemu_dim3 gdim3();
uint32_t lane();
uint64_t coll(int op, uint64_t x, uint64_t y);
void sync_block();
int sync_block_or(int x);
void nap();
uint64_t clock_ticks();
// weak-memory store buffer of the running work-item (emu_sched.h)
bool sb_put(void* p, uint64_t v, int size);        // false: store directly instead
bool sb_get(const void* p, int size, uint64_t* v);  // forwarded from the own buffer
void sb_touch(const void* p, int size);             // drain if an entry overlaps [p, p+size)
void sb_drain();
enum Op { OP_BALLOT = 1, OP_FIRSTLANE, OP_READLANE, OP_DPP, OP_SHFL_XOR, OP_WAVE_BARRIER };
}  // namespace kcemu

#define threadIdx (kcemu::tid3())
#define blockIdx (kcemu::bid3())
#define blockDim (kcemu::bdim3())
#define gridDim (kcemu::gdim3())

inline uint64_t __ballot(int pred) { return kcemu::coll(kcemu::OP_BALLOT, pred ? 1 : 0, 0); }
inline int __popcll(uint64_t x) { return __builtin_popcountll(x); }
inline int __shfl_xor(int v, int o) { return (int)kcemu::coll(kcemu::OP_SHFL_XOR, (uint32_t)v, (uint64_t)o); }
inline void __syncthreads() { kcemu::sync_block(); }
inline int __syncthreads_or(int x) { return kcemu::sync_block_or(x); }
inline uint64_t wall_clock64() { return kcemu::clock_ticks(); }

inline uint32_t __builtin_amdgcn_mbcnt_lo(uint32_t mask, uint32_t add) {
    const uint32_t l = kcemu::lane();
    const uint32_t below = l >= 32 ? mask : (mask & ((1u << l) - 1u));
    return add + (uint32_t)__builtin_popcount(below);
}
inline uint32_t __builtin_amdgcn_mbcnt_hi(uint32_t mask, uint32_t add) {
    const uint32_t l = kcemu::lane();
    const uint32_t below = l < 32 ? 0u : (l == 63 ? (mask & 0x7FFFFFFFu) : (mask & ((1u << (l - 32)) - 1u)));
    return add + (uint32_t)__builtin_popcount(below);
}
inline uint32_t __builtin_amdgcn_readfirstlane(uint32_t x) { return (uint32_t)kcemu::coll(kcemu::OP_FIRSTLANE, x, 0); }
inline int __builtin_amdgcn_readlane(int x, int l) { return (int)kcemu::coll(kcemu::OP_READLANE, (uint32_t)x, (uint64_t)l); }
inline int __builtin_amdgcn_update_dpp(int old, int v, int ctrl, int row_mask, int bank_mask, bool bound_ctrl) {
    (void)bank_mask;
    (void)bound_ctrl;
    const uint64_t packed = (uint64_t)(uint32_t)ctrl | ((uint64_t)(uint32_t)row_mask << 16) | ((uint64_t)(uint32_t)old << 32);
    return (int)kcemu::coll(kcemu::OP_DPP, (uint32_t)v, packed);
}
inline void __builtin_amdgcn_wave_barrier() { (void)kcemu::coll(kcemu::OP_WAVE_BARRIER, 0, 0); }
#define __builtin_amdgcn_s_sleep(n) kcemu::nap()
#define __builtin_amdgcn_s_waitcnt(n) (kcemu::sb_drain(), __atomic_thread_fence(__ATOMIC_SEQ_CST))
#define __builtin_amdgcn_fence(order, scope) (kcemu::sb_drain(), __atomic_thread_fence(__ATOMIC_SEQ_CST))

// buffer resources: base + size; loads past the size return zeros, as on the GPU
struct kcemu_rsrc {
    const char* base;
    uint32_t bytes;
};
typedef kcemu_rsrc __amdgpu_buffer_rsrc_t;
inline kcemu_rsrc __builtin_amdgcn_make_buffer_rsrc(void* p, short stride, int bytes, int flags) {
    (void)stride;
    (void)flags;
    return kcemu_rsrc{static_cast<const char*>(p), (uint32_t)bytes};
}
typedef unsigned int kcemu_u32x4 __attribute__((ext_vector_type(4)));
inline kcemu_u32x4 __builtin_amdgcn_raw_buffer_load_b128(kcemu_rsrc r, int off, int soff, int aux) {
    (void)soff;
    (void)aux;
    kcemu_u32x4 v = {0, 0, 0, 0};
    const uint64_t o = (uint32_t)off;
    if (o + 16 <= r.bytes) {
        const uint32_t* p = reinterpret_cast<const uint32_t*>(r.base + o);
        kcemu::sb_touch(p, 16);
        v.x = __atomic_load_n(p, __ATOMIC_RELAXED);
        v.y = __atomic_load_n(p + 1, __ATOMIC_RELAXED);
        v.z = __atomic_load_n(p + 2, __ATOMIC_RELAXED);
        v.w = __atomic_load_n(p + 3, __ATOMIC_RELAXED);
    }
    return v;
}

template <class T>
inline T kcemu_load(T* p) {
    uint64_t v;
    if (kcemu::sb_get(p, (int)sizeof(T), &v)) return (T)v;
    return __atomic_load_n(p, __ATOMIC_SEQ_CST);
}
template <class T, class U>
inline void kcemu_store(T* p, U v, int order) {
    const T x = (T)v;
    if (order == __ATOMIC_RELAXED && kcemu::sb_put(p, (uint64_t)x, (int)sizeof(T))) return;
    kcemu::sb_drain();  // a release store orders the work-item's earlier stores before it
    __atomic_store_n(p, x, __ATOMIC_SEQ_CST);
}
template <class T>
inline T* kcemu_rmw(T* p) {
    kcemu::sb_touch(p, (int)sizeof(T));  // an own buffered store to the address goes first
    return p;
}

#define __hip_atomic_load(p, order, scope) kcemu_load((p))
#define __hip_atomic_store(p, v, order, scope) kcemu_store((p), (v), (order))
#define __hip_atomic_fetch_add(p, v, order, scope) __atomic_fetch_add(kcemu_rmw(p), (v), __ATOMIC_SEQ_CST)
#define __hip_atomic_fetch_sub(p, v, order, scope) __atomic_fetch_sub(kcemu_rmw(p), (v), __ATOMIC_SEQ_CST)
#define __hip_atomic_fetch_or(p, v, order, scope) __atomic_fetch_or(kcemu_rmw(p), (v), __ATOMIC_SEQ_CST)
#define __hip_atomic_fetch_min(p, v, order, scope) kcemu_fetch_min(kcemu_rmw(p), (v))
#define __hip_atomic_fetch_max(p, v, order, scope) kcemu_fetch_max(kcemu_rmw(p), (v))
#define __hip_atomic_compare_exchange_strong(p, e, d, o1, o2, scope) \
    __atomic_compare_exchange_n(kcemu_rmw(p), (e), (d), false, __ATOMIC_SEQ_CST, __ATOMIC_SEQ_CST)

template <class T>
inline T kcemu_fetch_min(T* p, T v) {
    T cur = __atomic_load_n(p, __ATOMIC_SEQ_CST);
    while (v < cur && !__atomic_compare_exchange_n(p, &cur, v, false, __ATOMIC_SEQ_CST, __ATOMIC_SEQ_CST)) {
    }
    return cur;
}
template <class T>
inline T kcemu_fetch_max(T* p, T v) {
    T cur = __atomic_load_n(p, __ATOMIC_SEQ_CST);
    while (v > cur && !__atomic_compare_exchange_n(p, &cur, v, false, __ATOMIC_SEQ_CST, __ATOMIC_SEQ_CST)) {
    }
    return cur;
}
