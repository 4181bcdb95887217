// CPU stand-in for the HIP host runtime, used only by tests/hostemu to run this project's host code
// (kcore_hip.hip, serve.hip, tune.cpp — the C ABI of include/kcore.h) on the CPU, with every kernel
// executed by the wave64 emulator of tests/emu. Test infrastructure, not a HIP runtime.
//
// Streams are in-order queues, each drained by its own thread; a kernel launch enqueues an
// emulated launch of the kernel (grid clamped to KCEMU_MAX_GRID workgroups, default 6: every
// kernel here is grid-stride or sizes its work by the grid it gets); copies and fills run in
// stream order; events record completion times; stream capture records the enqueued work into a
// graph that a launch replays with the arguments captured. As in HIP, hipFree, hipHostFree and
// hipDeviceSynchronize wait for every stream to be idle, so a free while a resident kernel runs
// blocks — here it aborts with a message after KCEMU_FREE_WAIT_S seconds (default 20) instead of
// hanging. hipMalloc / hipHostMalloc fill new memory with 0xA5 bytes: nothing may rely on zeroed
// allocations. The device is one agent with KCEMU_CUS compute units (default 4) and a 100 MHz wall
// clock (the emulator's 10 ns ticks; see KCEMU_WALL_KHZ below).
#pragma once
#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <functional>
#include <map>
#include <mutex>
#include <thread>
#include <vector>

#include "../../emu/hip/hip_runtime.h"
#include "../../emu/emu_sched.h"

enum hipError_t {
    hipSuccess = 0,
    hipErrorInvalidValue = 1,
    hipErrorOutOfMemory = 2,
    hipErrorNoDevice = 100,
    hipErrorInvalidDevice = 101,
    hipErrorInvalidHandle = 400,
    hipErrorNotReady = 600,
    hipErrorUnknown = 999,
};
enum hipMemcpyKind {
    hipMemcpyHostToHost = 0,
    hipMemcpyHostToDevice = 1,
    hipMemcpyDeviceToHost = 2,
    hipMemcpyDeviceToDevice = 3,
    hipMemcpyDefault = 4,
};
enum hipMemoryType { hipMemoryTypeHost = 0, hipMemoryTypeDevice = 1 };
enum hipDeviceAttribute_t { hipDeviceAttributeWallClockRate = 1 };
enum hipStreamCaptureMode { hipStreamCaptureModeGlobal = 0, hipStreamCaptureModeThreadLocal, hipStreamCaptureModeRelaxed };
#define hipStreamNonBlocking 1u
#define hipEventDisableTiming 2u
#define hipHostMallocDefault 0u
#define hipHostMallocCoherent 0x40000000u

struct dim3 {
    uint32_t x, y, z;
    constexpr dim3(uint32_t a = 1, uint32_t b = 1, uint32_t c = 1) : x(a), y(b), z(c) {}
};
struct hipDeviceProp_t {
    char name[256];
    int multiProcessorCount;
    char gcnArchName[256];
};
struct hipPointerAttribute_t {
    hipMemoryType type;
    int device;
    void* devicePointer;
    void* hostPointer;
};

namespace kchost {

inline double now() {
    using namespace std::chrono;
    return duration<double>(steady_clock::now().time_since_epoch()).count();
}
inline uint32_t env_u32(const char* k, uint32_t d) {
    const char* e = std::getenv(k);
    return e && *e ? (uint32_t)std::strtoul(e, nullptr, 10) : d;
}

using Task = std::function<void()>;

}  // namespace kchost

struct ihipStream_t {
    std::mutex m;
    std::condition_variable cv;
    std::deque<kchost::Task> q;
    bool busy = false, quit = false;
    std::thread th;
    std::vector<kchost::Task>* capture = nullptr;
};
typedef ihipStream_t* hipStream_t;
struct ihipEvent_t {
    std::atomic<int> state{2};  // 0 recorded and pending, 1 complete, 2 never recorded
    double t = 0;
};
typedef ihipEvent_t* hipEvent_t;
struct ihipGraph {
    std::vector<kchost::Task> tasks;
};
typedef ihipGraph* hipGraph_t;
struct ihipGraphExec {
    std::vector<kchost::Task> tasks;
};
typedef ihipGraphExec* hipGraphExec_t;
typedef void* hipGraphNode_t;

namespace kchost {

struct Registry {
    std::mutex m;
    std::vector<hipStream_t> streams;
    std::map<uintptr_t, size_t> dev, host;  // base -> bytes
};
inline Registry& reg() {
    static Registry* r = new Registry;  // never destroyed: worker threads may outlive main
    return *r;
}
inline hipStream_t null_stream();

inline void worker(hipStream_t s) {
    std::unique_lock<std::mutex> l(s->m);
    while (true) {
        s->cv.wait(l, [&] { return s->quit || !s->q.empty(); });
        if (s->q.empty()) return;
        Task t = std::move(s->q.front());
        s->q.pop_front();
        s->busy = true;
        l.unlock();
        t();
        l.lock();
        s->busy = false;
        s->cv.notify_all();
    }
}
inline hipStream_t make_stream() {
    hipStream_t s = new ihipStream_t;
    s->th = std::thread(worker, s);
    std::lock_guard<std::mutex> l(reg().m);
    reg().streams.push_back(s);
    return s;
}
inline hipStream_t null_stream() {
    static hipStream_t s = make_stream();
    return s;
}
inline hipStream_t resolve(hipStream_t s) { return s ? s : null_stream(); }
inline void enqueue(hipStream_t s, Task t) {
    s = resolve(s);
    std::lock_guard<std::mutex> l(s->m);
    if (s->capture) {
        s->capture->push_back(std::move(t));
        return;
    }
    s->q.push_back(std::move(t));
    s->cv.notify_all();
}
inline bool idle(hipStream_t s) {
    std::lock_guard<std::mutex> l(s->m);
    return s->q.empty() && !s->busy;
}
inline void wait_idle(hipStream_t s) {
    std::unique_lock<std::mutex> l(s->m);
    s->cv.wait(l, [&] { return s->q.empty() && !s->busy; });
}
// every stream idle; `limit_s` > 0: abort when that takes longer (a free waiting on a resident kernel)
inline void device_sync(const char* who, double limit_s) {
    std::vector<hipStream_t> v;
    {
        std::lock_guard<std::mutex> l(reg().m);
        v = reg().streams;
    }
    const double t0 = now();
    for (hipStream_t s : v) {
        while (!idle(s)) {
            if (limit_s > 0 && now() - t0 > limit_s) {
                std::fprintf(stderr,
                             "hostemu: %s still waits for a running stream after %.0f s (it synchronises the "
                             "whole device: a kernel that does not end, such as a resident serving kernel)\n",
                             who, limit_s);
                std::fflush(stderr);
                std::abort();
            }
            std::this_thread::sleep_for(std::chrono::microseconds(200));
        }
    }
}
inline double free_wait() { return (double)env_u32("KCEMU_FREE_WAIT_S", 20); }
// KCEMU_EXACT=1: exactly the bytes asked for (with AddressSanitizer, any access past an
// allocation's end is reported); otherwise page granularity plus slack, as device memory has.
inline void* alloc_filled(size_t bytes, size_t align) {
    void* p = nullptr;
    const size_t n = env_u32("KCEMU_EXACT", 0) ? (bytes ? bytes : 1) : ((bytes + 4095) & ~(size_t)4095) + 4096;
    if (posix_memalign(&p, align, n)) return nullptr;
    std::memset(p, 0xA5, n);
    return p;
}
inline bool find(std::map<uintptr_t, size_t>& m, const void* p, uintptr_t* base) {
    const uintptr_t a = reinterpret_cast<uintptr_t>(p);
    auto it = m.upper_bound(a);
    if (it == m.begin()) return false;
    --it;
    if (a >= it->first + it->second) return false;
    *base = it->first;
    return true;
}

}  // namespace kchost

inline hipError_t hipGetLastError() { return hipSuccess; }
inline hipError_t hipSetDevice(int d) { return d == 0 ? hipSuccess : hipErrorInvalidDevice; }
inline hipError_t hipGetDeviceCount(int* n) {
    *n = 1;
    return hipSuccess;
}
inline hipError_t hipGetDeviceProperties(hipDeviceProp_t* p, int d) {
    if (d != 0) return hipErrorInvalidDevice;
    std::memset(p, 0, sizeof(*p));
    std::snprintf(p->name, sizeof(p->name), "wave64 emulator");
    std::snprintf(p->gcnArchName, sizeof(p->gcnArchName), "gfx950-emu");
    p->multiProcessorCount = (int)kchost::env_u32("KCEMU_CUS", 4);
    return hipSuccess;
}
inline hipError_t hipDeviceGetAttribute(int* v, hipDeviceAttribute_t a, int d) {
    if (d != 0) return hipErrorInvalidDevice;
    // KCEMU_WALL_KHZ scales every tick-based window (registration, wait limits, lease) against
    // the emulator's 10 ns clock: a larger value gives slowly starting OS threads time to join
    *v = a == hipDeviceAttributeWallClockRate ? (int)kchost::env_u32("KCEMU_WALL_KHZ", 100000) : 0;
    return hipSuccess;
}
template <class F>
inline hipError_t hipOccupancyMaxActiveBlocksPerMultiprocessor(int* n, F, int, size_t) {
    *n = 1;
    return hipSuccess;
}
inline hipError_t hipDeviceSynchronize() {
    kchost::device_sync("hipDeviceSynchronize", kchost::free_wait());
    return hipSuccess;
}

inline hipError_t hipStreamCreateWithFlags(hipStream_t* s, unsigned) {
    *s = kchost::make_stream();
    return hipSuccess;
}
inline hipError_t hipStreamSynchronize(hipStream_t s) {
    kchost::wait_idle(kchost::resolve(s));
    return hipSuccess;
}
inline hipError_t hipStreamQuery(hipStream_t s) { return kchost::idle(kchost::resolve(s)) ? hipSuccess : hipErrorNotReady; }
inline hipError_t hipStreamDestroy(hipStream_t s) {
    if (!s) return hipErrorInvalidHandle;
    kchost::wait_idle(s);
    {
        std::lock_guard<std::mutex> l(s->m);
        s->quit = true;
        s->cv.notify_all();
    }
    s->th.join();
    {
        std::lock_guard<std::mutex> l(kchost::reg().m);
        auto& v = kchost::reg().streams;
        v.erase(std::remove(v.begin(), v.end(), s), v.end());
    }
    delete s;
    return hipSuccess;
}
inline hipError_t hipStreamBeginCapture(hipStream_t s, hipStreamCaptureMode) {
    s = kchost::resolve(s);
    std::lock_guard<std::mutex> l(s->m);
    if (s->capture) return hipErrorInvalidValue;
    s->capture = new std::vector<kchost::Task>;
    return hipSuccess;
}
inline hipError_t hipStreamEndCapture(hipStream_t s, hipGraph_t* g) {
    s = kchost::resolve(s);
    std::lock_guard<std::mutex> l(s->m);
    if (!s->capture) return hipErrorInvalidValue;
    *g = new ihipGraph{std::move(*s->capture)};
    delete s->capture;
    s->capture = nullptr;
    return hipSuccess;
}
inline hipError_t hipGraphInstantiate(hipGraphExec_t* e, hipGraph_t g, hipGraphNode_t*, char*, size_t) {
    *e = new ihipGraphExec{g->tasks};
    return hipSuccess;
}
inline hipError_t hipGraphLaunch(hipGraphExec_t e, hipStream_t s) {
    kchost::enqueue(s, [e] {
        for (auto& t : e->tasks) t();
    });
    return hipSuccess;
}
inline hipError_t hipGraphDestroy(hipGraph_t g) {
    delete g;
    return hipSuccess;
}
inline hipError_t hipGraphExecDestroy(hipGraphExec_t e) {
    delete e;
    return hipSuccess;
}

inline hipError_t hipEventCreate(hipEvent_t* e) {
    *e = new ihipEvent_t;
    return hipSuccess;
}
inline hipError_t hipEventCreateWithFlags(hipEvent_t* e, unsigned) { return hipEventCreate(e); }
inline hipError_t hipEventDestroy(hipEvent_t e) {
    delete e;
    return hipSuccess;
}
inline hipError_t hipEventRecord(hipEvent_t e, hipStream_t s) {
    e->state = 0;
    kchost::enqueue(s, [e] {
        e->t = kchost::now();
        e->state = 1;
    });
    return hipSuccess;
}
inline hipError_t hipEventQuery(hipEvent_t e) { return e->state == 0 ? hipErrorNotReady : hipSuccess; }
inline hipError_t hipEventSynchronize(hipEvent_t e) {
    while (e->state == 0) std::this_thread::sleep_for(std::chrono::microseconds(100));
    return hipSuccess;
}
inline hipError_t hipEventElapsedTime(float* ms, hipEvent_t a, hipEvent_t b) {
    if (a->state != 1 || b->state != 1) return hipErrorNotReady;
    *ms = (float)((b->t - a->t) * 1e3);
    return hipSuccess;
}

inline hipError_t hipMalloc(void** p, size_t bytes) {
    *p = kchost::alloc_filled(bytes, 256);
    if (!*p) return hipErrorOutOfMemory;
    std::lock_guard<std::mutex> l(kchost::reg().m);
    kchost::reg().dev[reinterpret_cast<uintptr_t>(*p)] = bytes ? bytes : 1;
    return hipSuccess;
}
template <class T>
inline hipError_t hipMalloc(T** p, size_t bytes) {
    return hipMalloc(reinterpret_cast<void**>(p), bytes);
}
inline hipError_t hipFree(void* p) {
    if (!p) return hipSuccess;
    kchost::device_sync("hipFree", kchost::free_wait());
    {
        std::lock_guard<std::mutex> l(kchost::reg().m);
        if (!kchost::reg().dev.erase(reinterpret_cast<uintptr_t>(p))) return hipErrorInvalidValue;
    }
    std::free(p);
    return hipSuccess;
}
inline hipError_t hipHostMalloc(void** p, size_t bytes, unsigned) {
    *p = kchost::alloc_filled(bytes, 4096);
    if (!*p) return hipErrorOutOfMemory;
    std::lock_guard<std::mutex> l(kchost::reg().m);
    kchost::reg().host[reinterpret_cast<uintptr_t>(*p)] = bytes ? bytes : 1;
    return hipSuccess;
}
template <class T>
inline hipError_t hipHostMalloc(T** p, size_t bytes, unsigned flags) {
    return hipHostMalloc(reinterpret_cast<void**>(p), bytes, flags);
}
inline hipError_t hipHostFree(void* p) {
    if (!p) return hipSuccess;
    kchost::device_sync("hipHostFree", kchost::free_wait());
    {
        std::lock_guard<std::mutex> l(kchost::reg().m);
        if (!kchost::reg().host.erase(reinterpret_cast<uintptr_t>(p))) return hipErrorInvalidValue;
    }
    std::free(p);
    return hipSuccess;
}
inline hipError_t hipPointerGetAttributes(hipPointerAttribute_t* at, const void* p) {
    std::lock_guard<std::mutex> l(kchost::reg().m);
    uintptr_t b = 0;
    std::memset(at, 0, sizeof(*at));
    if (kchost::find(kchost::reg().host, p, &b)) {
        at->type = hipMemoryTypeHost;
        at->hostPointer = const_cast<void*>(p);
        at->devicePointer = const_cast<void*>(p);
        return hipSuccess;
    }
    if (kchost::find(kchost::reg().dev, p, &b)) {
        at->type = hipMemoryTypeDevice;
        at->devicePointer = const_cast<void*>(p);
        return hipSuccess;
    }
    return hipErrorInvalidValue;
}
inline hipError_t hipMemcpyAsync(void* dst, const void* src, size_t bytes, hipMemcpyKind, hipStream_t s) {
    if (bytes) kchost::enqueue(s, [=] { std::memcpy(dst, src, bytes); });
    return hipSuccess;
}
inline hipError_t hipMemsetAsync(void* dst, int v, size_t bytes, hipStream_t s) {
    if (bytes) kchost::enqueue(s, [=] { std::memset(dst, v, bytes); });
    return hipSuccess;
}

// kernel<<<grid, block, shmem, stream>>>(args) is rewritten by the test build as
// kcemu_launch(grid, block, stream, kernel, args): the arguments are copied at the launch
template <class F, class... A>
inline void kcemu_launch(dim3 grid, dim3 block, hipStream_t s, F f, A... a) {
    const uint32_t g = std::max<uint32_t>(1, std::min<uint32_t>(grid.x, kchost::env_u32("KCEMU_MAX_GRID", 6)));
    const uint32_t b = block.x;
    kchost::enqueue(s, [=] { kcemu::launch(g, b, [=] { f(a...); }); });
}
