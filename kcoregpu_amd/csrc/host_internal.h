// host_internal.h — state shared by the host-side translation units of libkcore_hip.so.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "../../include/kcore.h"

struct kc_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr, evk0 = nullptr, evk1 = nullptr;
    int cus = 0;
    uint64_t wall_rate_khz = 100000;
    uint64_t wait_limit_ms = 20000;  // one device-side wait; KCORE_WAIT_LIMIT_MS overrides
    char arch[64] = {0};             // gcnArchName, part of the tuning cache key
};

inline kc_status hip_status(hipError_t e) {
    if (e == hipSuccess) return KC_OK;
    if (e == hipErrorOutOfMemory) return KC_ERR_OOM;
    if (e == hipErrorNoDevice || e == hipErrorInvalidDevice) return KC_ERR_NO_DEVICE;
    return KC_ERR_HIP;
}
#define KC_TRY(x)                                    \
    do {                                             \
        hipError_t e_ = (x);                         \
        if (e_ != hipSuccess) return hip_status(e_); \
    } while (0)

// hipFree and hipHostFree synchronise the whole device, so while a serving kernel is resident they
// would wait for it forever. Every free in this library goes through these two functions: they free
// at once when no serving kernel runs, and otherwise keep the block until the last one has stopped
// (pinned blocks from kc_host_alloc are handed out again meanwhile). Thread-safe. (serve.hip)
void lib_free_device(void* p);
void lib_free_pinned(void* p);

// A graph whose device arrays live in caller-provided memory (no hipMalloc; kc_graph_free releases
// nothing): the serving lanes place large requests in the arena this way. graph_bytes is the size
// of that memory for V vertices, `arcs` arcs and maximum degree dmax. (kcore_hip.hip)
struct kc_graph;
uint64_t graph_bytes(uint32_t V, uint64_t arcs, uint32_t dmax);
kc_status graph_upload_into(kc_ctx* c, const kc_csr* h, void* mem, uint64_t bytes, kc_graph** out);

