// serve_kernel.h — the persistent serving kernel: continuous batching of k-core requests.
//
// The kernel stays resident while the server is open. Requests are published in a submission
// ring in device memory: the host copies a request's CSR into its arena region and then the ring
// entry, in that order, on one copy stream, so an entry is visible only after its data. Every
// workgroup claims the next ring index from a device counter, waits for that entry, peels the
// request with the per-graph workgroup peel shared with batched peeling (wg_peel_graph), and
// signals completion in pinned host memory. A request that arrives while others run is taken by
// the next idle workgroup at once: requests join and leave the running kernel individually,
// with no relaunch and no batch boundary.
//
// Coherence: a request's inputs are written by the copy engine while the kernel runs, possibly
// into memory a previous request used, so the claiming workgroup reads the entry with a
// system-scope acquire (which invalidates the caches it reads through) before touching the data,
// and publishes completion with a system-scope release (which writes back what it wrote) so the
// host's download of the core numbers sees the final values.
#pragma once
#include "batch_kernel.h"

namespace kcdev {

// submission ring entry (device memory, written by the host's copy stream)
struct ServeEntry {
    uint64_t seq;   // ring index + 1 once published
    uint64_t base;  // byte offset of the request's region in the arena
    uint32_t V;
    uint32_t arcs;
    uint32_t flags;
    uint32_t pad;
};

// completion record (pinned host memory, written by the kernel)
struct ServeDone {
    uint64_t seq;  // ring index + 1 once finished
    uint32_t kmax, levels, rounds, err;
    uint64_t pad;
};

// Arena region of a request: offsets (V+1), neighbours (arcs), input degrees, working degrees,
// peel order, each starting on a 256-byte boundary.
__host__ __device__ inline uint64_t serve_align(uint64_t x) { return (x + 255) & ~255ull; }
struct ServeRegion {
    uint64_t off, nbr, deg0, core, order, bytes;
};
__host__ __device__ inline ServeRegion serve_region(uint32_t V, uint32_t arcs) {
    ServeRegion r;
    r.off = 0;
    r.nbr = serve_align(((uint64_t)V + 1) * 4);
    r.deg0 = r.nbr + serve_align((uint64_t)arcs * 4);
    r.core = r.deg0 + serve_align((uint64_t)V * 4);
    r.order = r.core + serve_align((uint64_t)V * 4);
    r.bytes = r.order + serve_align((uint64_t)V * 4);
    return r;
}

struct ServeArgs {
    uint8_t* arena;
    ServeEntry* ring;      // device memory
    ServeDone* done;       // pinned host memory
    uint64_t* head;        // device claim counter
    const uint32_t* stop;  // device memory, written by the host's copy stream: non-zero = exit
    uint32_t slots;
};

__device__ __forceinline__ uint64_t ld_sys64(const uint64_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}
__device__ __forceinline__ uint32_t ld_sys(const uint32_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}

template <int BLOCK>
__global__ __launch_bounds__(BLOCK) void serve_kernel(ServeArgs a) {
    __shared__ BatchState<BLOCK> S;
    __shared__ ServeEntry E;
    __shared__ uint32_t s_slot, s_stop;
    while (true) {
        if (threadIdx.x == 0) {
            const uint64_t i = __hip_atomic_fetch_add(a.head, 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            const uint32_t slot = (uint32_t)(i % a.slots);
            uint32_t stop = 0, spins = 0;
            // an idle server waits here until the next request or the stop flag; the polled words
            // are written by the copy engine, read with system-scope loads, and every 256 polls an
            // acquire fence also drops any cached copy of them
            while (ld_sys64(&a.ring[slot].seq) != i + 1) {
                if (ld_sys(a.stop)) {
                    stop = 1;
                    break;
                }
                if (spins < 64) {
                    __builtin_amdgcn_s_sleep(4);
                } else {
                    __builtin_amdgcn_s_sleep(64);
                }
                if ((++spins & 255) == 0) __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "");
            }
            if (!stop) {
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "");  // system scope: see the copy engine's data
                E = a.ring[slot];
            }
            s_slot = slot;
            s_stop = stop;
        }
        __syncthreads();
        if (s_stop) break;
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
        const uint32_t V = E.V, arcs = E.arcs, flags = E.flags;
        const ServeRegion r = serve_region(V, arcs);
        uint8_t* base = a.arena + E.base;
        const uint32_t* off = reinterpret_cast<const uint32_t*>(base + r.off);
        const uint32_t* nbr = reinterpret_cast<const uint32_t*>(base + r.nbr);
        const uint32_t* deg0 = reinterpret_cast<const uint32_t*>(base + r.deg0);
        uint32_t* core = reinterpret_cast<uint32_t*>(base + r.core);
        uint32_t* order = reinterpret_cast<uint32_t*>(base + r.order);
        // working degrees: SC1 stores reach the device coherence point; the caches this workgroup
        // reads through were invalidated by the acquire above, so plain prechecks can only see
        // these values or later (lower) ones
        for (uint32_t v = threadIdx.x; v < V; v += BLOCK) st_rlx(core + v, deg0[v]);
        drain_memory();
        __syncthreads();
        uint32_t levels = 0, rounds = 0, err = 0;
        const uint32_t km = wg_peel_graph<BLOCK>(S, off, nbr, core, order, V, flags, &levels, &rounds, &err);
        drain_memory();
        __syncthreads();
        if (threadIdx.x == 0) {
            ServeDone* d = a.done + s_slot;
            d->kmax = km;
            d->levels = levels;
            d->rounds = rounds;
            d->err = err;
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "");  // system scope: results reach memory first
            __hip_atomic_store(&d->seq, E.seq, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
        }
        __syncthreads();
    }
}

}  // namespace kcdev
