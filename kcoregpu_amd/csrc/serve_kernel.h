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
    uint32_t* out;  // pinned host buffer for the core numbers (written by the kernel), or null
    uint64_t pad2;
};

// completion record (pinned host memory, written by the kernel)
enum : uint32_t { SERVE_OK = 0, SERVE_NO_PROGRESS = 1, SERVE_BAD_CSR = 2 };
struct ServeDone {
    uint64_t seq;  // ring index + 1 once finished
    uint32_t kmax, levels, rounds, err;
    uint64_t pad;
};

// Arena region of a request: offsets (V+1), neighbours (arcs), working degrees, peel order,
// each starting on a 256-byte boundary. Degrees are taken from the offsets on the device.
__host__ __device__ inline uint64_t serve_align(uint64_t x) { return (x + 255) & ~255ull; }
struct ServeRegion {
    uint64_t off, nbr, core, order, bytes;
};
__host__ __device__ inline ServeRegion serve_region(uint32_t V, uint32_t arcs) {
    ServeRegion r;
    r.off = 0;
    r.nbr = serve_align(((uint64_t)V + 1) * 4);
    r.core = r.nbr + serve_align((uint64_t)arcs * 4);
    r.order = r.core + serve_align((uint64_t)V * 4);
    r.bytes = r.order + serve_align((uint64_t)V * 4);
    return r;
}

struct ServeArgs {
    uint8_t* arena;
    ServeEntry* ring;       // device memory
    ServeDone* done;        // pinned host memory
    uint64_t* head;         // device claim counter
    const uint32_t* stop;   // device memory, written by the host's copy stream: non-zero = exit
    const uint32_t* hstop;  // pinned host memory, written by the host directly: non-zero = exit
    const uint32_t* hbeat;  // pinned host memory, advanced by the host's heartbeat thread
    uint32_t* expired;      // device word: set when a workgroup left because the lease ran out
    uint64_t lease_ticks;   // wall-clock ticks without a heartbeat after which an idle kernel exits
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
            // An idle server waits here until the next request or a stop flag. The polled words
            // are written by the copy engine, read with system-scope loads, and every 256 polls
            // an acquire fence also drops any cached copy of them. The host's heartbeat is a
            // lease: if the host process is gone (the heartbeat has stopped for lease_ticks),
            // the workgroup exits, so a dead host never leaves the kernel running.
            uint32_t beat = ld_sys(a.hbeat);
            uint64_t t_beat = wall_clock64();
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
                if ((++spins & 255) == 0) {
                    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "");
                    if (ld_sys(a.hstop)) {
                        stop = 1;
                        break;
                    }
                    const uint32_t b = ld_sys(a.hbeat);
                    const uint64_t now = wall_clock64();
                    if (b != beat) {
                        beat = b;
                        t_beat = now;
                    } else if (now - t_beat > a.lease_ticks) {
                        st_rlx(a.expired, 1u);
                        stop = 1;
                        break;
                    }
                }
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
        uint32_t* core = reinterpret_cast<uint32_t*>(base + r.core);
        uint32_t* order = reinterpret_cast<uint32_t*>(base + r.order);
        // The request is checked before anything indexes with it: offsets start at 0, end at
        // arcs and never decrease (so every adjacency lies inside the neighbour array), and
        // every neighbour id is below V. Working degrees come from the offsets; they are written
        // with SC1 stores that reach the device coherence point, and the caches this workgroup
        // reads through were invalidated by the acquire above, so plain prechecks can only see
        // these values or later (lower) ones.
        uint32_t bad = (threadIdx.x == 0 && (off[0] != 0 || off[V] != arcs)) ? 1u : 0u;
        for (uint32_t v = threadIdx.x; v < V; v += BLOCK) {
            const uint32_t o0 = off[v], o1 = off[v + 1];
            bad |= o1 < o0 ? 1u : 0u;
            st_rlx(core + v, o1 - o0);
        }
        for (uint32_t j = threadIdx.x; j < arcs; j += BLOCK) bad |= nbr[j] >= V ? 1u : 0u;
        drain_memory();
        bad = __syncthreads_or(bad);
        uint32_t levels = 0, rounds = 0, err = 0, km = 0;
        if (!bad) km = wg_peel_graph<BLOCK>(S, off, nbr, core, order, V, flags, &levels, &rounds, &err);
        drain_memory();
        __syncthreads();
        // results straight into the caller's pinned buffer when it gave one
        if (E.out && !bad)
            for (uint32_t v = threadIdx.x; v < V; v += BLOCK) E.out[v] = ld_rlx(core + v);
        drain_memory();
        __syncthreads();
        if (threadIdx.x == 0) {
            ServeDone* d = a.done + s_slot;
            d->kmax = km;
            d->levels = levels;
            d->rounds = rounds;
            d->err = bad ? SERVE_BAD_CSR : (err ? SERVE_NO_PROGRESS : SERVE_OK);
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "");  // system scope: results reach memory first
            __hip_atomic_store(&d->seq, E.seq, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
        }
        __syncthreads();
    }
}

}  // namespace kcdev
