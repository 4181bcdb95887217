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

