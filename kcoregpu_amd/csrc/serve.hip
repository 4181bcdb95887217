// serve.hip — host side of served peeling (kc_serve_* in include/kcore.h).
//
// A server owns: a device arena of `hbm_budget_bytes` from which each admitted request gets one
// region (first fit, 256-byte aligned; a request's footprint is its CSR plus three vertex arrays,
// about 8 m + 16 n bytes for n vertices and m edges); a submission ring in device memory; pinned
// completion records and a stop flag; two streams — the persistent serving kernel on one, the
// request uploads, ring publications and result downloads on the other, so copies overlap the
// peeling of earlier requests.
//
// Requests are admitted in arrival order whenever a ring slot is free and the arena has room
// (a request that does not fit waits; one larger than the whole arena is refused). kc_serve_poll
// moves completed requests' core numbers to the caller's buffers and frees their regions.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstring>
#include <deque>
#include <map>
#include <new>
#include <vector>

#include "host_internal.h"
#include "serve_kernel.h"

using namespace kcdev;

namespace {

double now_s() {
    using namespace std::chrono;
    return duration<double>(steady_clock::now().time_since_epoch()).count();
}

struct Req {
    uint64_t id = 0;
    uint32_t V = 0, arcs = 0;
    const uint32_t* off = nullptr;
    const uint32_t* nbr = nullptr;
    const uint32_t* deg = nullptr;
    uint32_t* core_out = nullptr;
    uint64_t base = 0;
    ServeRegion r{};
    uint64_t index = 0;
    double t_submit = 0, t_admit = 0;
    hipEvent_t ev = nullptr;
    int state = 0;  // 0 free, 1 on the device, 2 downloading
};

// first-fit allocator over [0, size)
struct Arena {
    std::map<uint64_t, uint64_t> free_;  // offset -> length
    uint64_t used = 0;
    void init(uint64_t size) {
        free_.clear();
        free_[0] = size;
        used = 0;
    }
    bool alloc(uint64_t n, uint64_t* off) {
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second < n) continue;
            *off = it->first;
            const uint64_t rest = it->second - n;
            const uint64_t at = it->first + n;
            free_.erase(it);
            if (rest) free_[at] = rest;
            used += n;
            return true;
        }
        return false;
    }
    void release(uint64_t off, uint64_t n) {
        used -= n;
        auto it = free_.emplace(off, n).first;
        auto nx = std::next(it);
        if (nx != free_.end() && it->first + it->second == nx->first) {
            it->second += nx->second;
            free_.erase(nx);
        }
        if (it != free_.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) {
                pv->second += it->second;
                free_.erase(it);
            }
        }
    }
};

template <int BLOCK>
hipError_t launch_serve(const ServeArgs& a, uint32_t grid, hipStream_t s) {
    serve_kernel<BLOCK><<<dim3(grid), dim3(BLOCK), 0, s>>>(a);
    return hipGetLastError();
}

}  // namespace

struct kc_server {
    kc_ctx* ctx = nullptr;
    kc_serve_cfg cfg{};
    hipStream_t kstream = nullptr, cstream = nullptr;
    uint8_t* arena = nullptr;
    ServeEntry* d_ring = nullptr;
    uint32_t* d_head = nullptr;  // words 0-1: claim counter (uint64); word 32: stop flag
    ServeDone* h_done = nullptr;
    ServeEntry* h_entry = nullptr;
    uint32_t* h_stop = nullptr;  // pinned source of the stop flag's value
    Arena alloc;
    std::deque<Req> queue;
    std::vector<Req> slot;
    uint64_t next_index = 0;
    uint32_t inflight = 0;
    bool running = false;
};

static constexpr uint32_t kStopWord = 32;

// Stops the serving kernel (the stop flag travels on the copy stream, after every publication)
// and frees everything. Returns KC_ERR_DEVICE when the kernel did not exit within the wait limit;
// its memory is then left allocated, since the kernel may still read it.
static kc_status serve_release(kc_server* s) {
    if (s->running) {
        *s->h_stop = 1;
        hipError_t e = hipMemcpyAsync(s->d_head + kStopWord, s->h_stop, 4, hipMemcpyHostToDevice, s->cstream);
        if (e) return hip_status(e);
        const double t0 = now_s();
        while ((e = hipStreamQuery(s->kstream)) == hipErrorNotReady) {
            if (now_s() - t0 > s->ctx->wait_limit_ms * 1e-3) return KC_ERR_DEVICE;
        }
        s->running = false;
    }
    if (s->cstream) (void)hipStreamSynchronize(s->cstream);
    for (Req& r : s->slot)
        if (r.ev) (void)hipEventDestroy(r.ev);
    for (void* p : {(void*)s->arena, (void*)s->d_ring, (void*)s->d_head})
        if (p) (void)hipFree(p);
    for (void* p : {(void*)s->h_done, (void*)s->h_entry, (void*)s->h_stop})
        if (p) (void)hipHostFree(p);
    if (s->kstream) (void)hipStreamDestroy(s->kstream);
    if (s->cstream) (void)hipStreamDestroy(s->cstream);
    delete s;
    return KC_OK;
}

// admit queued requests in arrival order while a slot is free and the arena has room
static kc_status serve_admit(kc_server* s) {
    while (!s->queue.empty()) {
        // ring slots are reused in index order: the next index waits for its slot's last occupant
        const uint32_t sl = (uint32_t)(s->next_index % s->cfg.ring_slots);
        Req& r = s->slot[sl];
        if (r.state != 0) break;
        Req& q = s->queue.front();
        uint64_t base;
        if (!s->alloc.alloc(q.r.bytes, &base)) break;
        hipEvent_t ev = r.ev;
        r = q;
        r.ev = ev;
        s->queue.pop_front();
        r.base = base;
        r.index = s->next_index++;
        r.t_admit = now_s();
        r.state = 1;
        s->inflight++;
        uint8_t* dst = s->arena + base;
        hipStream_t st = s->cstream;
        KC_TRY(hipMemcpyAsync(dst + r.r.off, r.off, ((size_t)r.V + 1) * 4, hipMemcpyHostToDevice, st));
        if (r.arcs) KC_TRY(hipMemcpyAsync(dst + r.r.nbr, r.nbr, (size_t)r.arcs * 4, hipMemcpyHostToDevice, st));
        KC_TRY(hipMemcpyAsync(dst + r.r.deg0, r.deg, (size_t)r.V * 4, hipMemcpyHostToDevice, st));
        ServeEntry* e = s->h_entry + sl;
        e->base = base;
        e->V = r.V;
        e->arcs = r.arcs;
        e->flags = s->cfg.flags;
        e->pad = 0;
        e->seq = r.index + 1;
        // the entry goes last on the same stream: the kernel sees it only after the data
        KC_TRY(hipMemcpyAsync(s->d_ring + sl, e, sizeof(ServeEntry), hipMemcpyHostToDevice, st));
    }
    return KC_OK;
}

extern "C" {

double kc_serve_clock(void) { return now_s(); }

kc_status kc_serve_open(kc_ctx* c, const kc_serve_cfg* cfg, kc_server** out) {
    if (!c || !out) return KC_ERR_ARG;
    KC_TRY(hipSetDevice(c->device));
    kc_server* s = new (std::nothrow) kc_server;
    if (!s) return KC_ERR_OOM;
    s->ctx = c;
    if (cfg) s->cfg = *cfg;
    if (!s->cfg.hbm_budget_bytes) s->cfg.hbm_budget_bytes = 8ull << 30;
    if (!s->cfg.ring_slots) s->cfg.ring_slots = 1024;
    if (!s->cfg.block) s->cfg.block = 256;
    if (s->cfg.block != 256 && s->cfg.block != 512 && s->cfg.block != 1024) {
        delete s;
        return KC_ERR_ARG;
    }
    // leave room on every CU for the runtime's own copy kernels next to the resident server
    if (!s->cfg.workgroups) s->cfg.workgroups = (uint32_t)c->cus * (s->cfg.block == 256 ? 2 : 1);
    const uint32_t R = s->cfg.ring_slots;
    auto fail = [&](hipError_t e) {
        (void)serve_release(s);
        return hip_status(e);
    };
    hipError_t e;
    if ((e = hipStreamCreateWithFlags(&s->kstream, hipStreamNonBlocking))) return fail(e);
    if ((e = hipStreamCreateWithFlags(&s->cstream, hipStreamNonBlocking))) return fail(e);
    if ((e = hipMalloc(&s->arena, s->cfg.hbm_budget_bytes))) return fail(e);
    if ((e = hipMalloc(&s->d_ring, (size_t)R * sizeof(ServeEntry)))) return fail(e);
    if ((e = hipMalloc(&s->d_head, 256))) return fail(e);
    // completion records are written by the kernel and polled by the host: coherent memory
    if ((e = hipHostMalloc(&s->h_done, (size_t)R * sizeof(ServeDone), hipHostMallocCoherent))) return fail(e);
    if ((e = hipHostMalloc(&s->h_entry, (size_t)R * sizeof(ServeEntry), hipHostMallocDefault))) return fail(e);
    if ((e = hipHostMalloc(&s->h_stop, 256, hipHostMallocDefault))) return fail(e);
    memset(s->h_done, 0, (size_t)R * sizeof(ServeDone));
    memset(s->h_entry, 0, (size_t)R * sizeof(ServeEntry));
    *s->h_stop = 0;
    if ((e = hipMemsetAsync(s->d_ring, 0, (size_t)R * sizeof(ServeEntry), s->kstream))) return fail(e);
    if ((e = hipMemsetAsync(s->d_head, 0, 256, s->kstream))) return fail(e);
    // the ring must be clear before the copy stream can publish into it
    if ((e = hipStreamSynchronize(s->kstream))) return fail(e);
    s->slot.resize(R);
    for (Req& r : s->slot)
        if ((e = hipEventCreateWithFlags(&r.ev, hipEventDisableTiming))) return fail(e);
    s->alloc.init(s->cfg.hbm_budget_bytes);
    ServeArgs a{};
    a.arena = s->arena;
    a.ring = s->d_ring;
    a.done = s->h_done;
    a.head = reinterpret_cast<uint64_t*>(s->d_head);
    a.stop = s->d_head + kStopWord;
    a.slots = R;
    switch (s->cfg.block) {
        case 256: e = launch_serve<256>(a, s->cfg.workgroups, s->kstream); break;
        case 512: e = launch_serve<512>(a, s->cfg.workgroups, s->kstream); break;
        default: e = launch_serve<1024>(a, s->cfg.workgroups, s->kstream); break;
    }
    if (e) return fail(e);
    s->running = true;
    *out = s;
    return KC_OK;
}

kc_status kc_serve_submit(kc_server* s, uint64_t req_id, const kc_csr* g, uint32_t* core_out) {
    if (!s || !g || !g->offsets || (g->arcs && !g->neighbors) || !g->degrees) return KC_ERR_ARG;
    if (g->V == 0 || g->V >= (1u << 30) || g->arcs >= 0xFFFFFFFFull) return KC_ERR_CAPACITY;
    if (g->offsets[0] != 0 || g->offsets[g->V] != g->arcs) return KC_ERR_ARG;
    Req q;
    q.id = req_id;
    q.V = g->V;
    q.arcs = (uint32_t)g->arcs;
    q.off = g->offsets;
    q.nbr = g->neighbors;
    q.deg = g->degrees;
    q.core_out = core_out;
    q.r = serve_region(q.V, q.arcs);
    if (q.r.bytes > s->cfg.hbm_budget_bytes) return KC_ERR_CAPACITY;
    q.t_submit = now_s();
    s->queue.push_back(q);
    return serve_admit(s);
}

kc_status kc_serve_poll(kc_server* s, kc_completion* out, uint32_t cap, uint32_t* n_out) {
    if (!s || !n_out || (cap && !out)) return KC_ERR_ARG;
    uint32_t n = 0;
    for (uint32_t sl = 0; sl < s->cfg.ring_slots && s->inflight; ++sl) {
        Req& r = s->slot[sl];
        if (r.state == 1) {
            const uint64_t seq = __atomic_load_n(&s->h_done[sl].seq, __ATOMIC_ACQUIRE);
            if (seq != r.index + 1) continue;
            if (r.core_out) {
                KC_TRY(hipMemcpyAsync(r.core_out, s->arena + r.base + r.r.core, (size_t)r.V * 4,
                                      hipMemcpyDeviceToHost, s->cstream));
            }
            KC_TRY(hipEventRecord(r.ev, s->cstream));
            r.state = 2;
        }
        if (r.state == 2 && n < cap) {
            const hipError_t q = hipEventQuery(r.ev);
            if (q == hipErrorNotReady) continue;
            KC_TRY(q);
            kc_completion& c = out[n++];
            const ServeDone& d = s->h_done[sl];
            c.req_id = r.id;
            c.kmax = d.kmax;
            c.levels = d.levels;
            c.rounds = d.rounds;
            c.status = d.err ? KC_ERR_NO_PROGRESS : KC_OK;
            c.t_submit = r.t_submit;
            c.t_admit = r.t_admit;
            c.t_done = now_s();
            s->alloc.release(r.base, r.r.bytes);
            r.state = 0;
            s->inflight--;
        }
    }
    *n_out = n;
    return serve_admit(s);
}

kc_status kc_serve_pending(const kc_server* s, uint32_t* queued, uint32_t* inflight, uint64_t* arena_used) {
    if (!s) return KC_ERR_ARG;
    if (queued) *queued = (uint32_t)s->queue.size();
    if (inflight) *inflight = s->inflight;
    if (arena_used) *arena_used = s->alloc.used;
    return KC_OK;
}

kc_status kc_serve_close(kc_server* s) {
    if (!s) return KC_ERR_ARG;
    KC_TRY(hipSetDevice(s->ctx->device));
    return serve_release(s);
}

kc_status kc_host_alloc(uint64_t bytes, void** out) {
    if (!out) return KC_ERR_ARG;
    *out = nullptr;
    if (!bytes) return KC_OK;
    KC_TRY(hipHostMalloc(out, bytes, hipHostMallocDefault));
    return KC_OK;
}

kc_status kc_host_free(void* p) {
    if (p) KC_TRY(hipHostFree(p));
    return KC_OK;
}

}  // extern "C"
