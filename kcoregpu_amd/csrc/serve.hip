// serve.hip — host side of served peeling (kc_serve_* in include/kcore.h).
//
// A server owns: a device arena of `hbm_budget_bytes` from which each admitted request gets one
// region (first fit, 256-byte aligned: offsets, neighbours, working degrees, peel order); a
// submission ring in device memory; pinned completion records; a pinned control page (stop flag
// and heartbeat); two streams — the persistent serving kernel on one, request uploads and ring
// publications on the other, so copies overlap the peeling of earlier requests.
//
// Requests are admitted in arrival order whenever their ring slot is free and the arena has room
// for their charge, max(8 m + 16 n bytes, region size) for n vertices and m edges; a request that
// does not fit waits, one larger than the whole arena is refused. A request whose core_out is
// pinned host memory gets its core numbers written there by the kernel itself; otherwise
// kc_serve_poll copies them down.
//
// Large requests (at least large_arcs arcs) would hold one workgroup for long: they go to a
// large-request lane instead, where a worker thread with its own context and stream places the
// request's whole single-graph state in its arena region, peels it with the single-graph kernel
// over many workgroups (running beside the resident kernel) and downloads the core numbers;
// large_lanes workers run at once. Lane contexts are opened before the resident kernel starts and
// closed after it ends, and a lane allocates and frees no device memory.
//
// Lifetime safety: the kernel never outlives its process. A heartbeat thread advances a word in
// the control page while the server is open; idle kernel workgroups exit when it stops advancing
// for lease_ms, and an exit handler stops every open server before the HIP runtime shuts down.
//
// hipFree and hipHostFree synchronise the whole device, so while a serving kernel is resident
// they would wait for it forever: the library's frees (lib_free_device, lib_free_pinned,
// kc_host_free) go through one FreeRegistry (resident.h), which defers them until the last
// resident kernel has ended and hands returned pinned blocks out again meanwhile.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <new>
#include <thread>
#include <vector>

#include "arena.h"
#include "host_internal.h"
#include "resident.h"
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
    uint32_t* core_out = nullptr;
    bool out_pinned = false;
    uint64_t base = 0, charge = 0;
    ServeRegion r{};
    uint64_t index = 0;
    double t_submit = 0, t_admit = 0;
    hipEvent_t ev = nullptr;
    int state = 0;  // 0 free, 1 on the device, 2 finishing
    bool large = false;
};

template <int BLOCK>
hipError_t launch_serve(const ServeArgs& a, uint32_t grid, hipStream_t s) {
    serve_kernel<BLOCK><<<dim3(grid), dim3(BLOCK), 0, s>>>(a);
    return hipGetLastError();
}

constexpr uint32_t kStopWord = 32;  // in d_head: device stop flag (stream-ordered)
constexpr uint32_t kExpired = 48;   // in d_head: set by a workgroup that left on lease expiry
constexpr uint32_t kHStop = 0;      // in the pinned control page: immediate stop flag
constexpr uint32_t kHBeat = 16;     // in the pinned control page: heartbeat counter

bool is_pinned(const void* p) {
    hipPointerAttribute_t at;
    if (hipPointerGetAttributes(&at, p) != hipSuccess) {
        (void)hipGetLastError();
        return false;
    }
    return at.type == hipMemoryTypeHost && at.hostPointer != nullptr;
}

}  // namespace

struct kc_server {
    kc_ctx* ctx = nullptr;
    kc_serve_cfg cfg{};
    hipStream_t kstream = nullptr, cstream = nullptr;
    uint8_t* arena = nullptr;
    ServeEntry* d_ring = nullptr;
    uint32_t* d_head = nullptr;  // words 0-1: claim counter (uint64); kStopWord; kExpired
    ServeDone* h_done = nullptr;
    ServeEntry* h_entry = nullptr;
    uint32_t* h_ctl = nullptr;  // pinned, coherent: kHStop, kHBeat
    uint32_t* h_one = nullptr;  // pinned source of the device stop flag's value
    Arena alloc;
    std::deque<Req> queue;
    std::vector<Req> slot;
    std::vector<uint32_t> fresh;  // slots admitted by one serve_admit call, to publish
    uint64_t next_index = 0;
    uint32_t inflight = 0;
    bool running = false;
    bool resident = false;  // counted in the resident-kernel registry (launched, end not yet seen)
    std::thread hb;  // heartbeat
    std::mutex hb_m;
    std::condition_variable hb_cv;
    bool hb_quit = false;
    // large-request lane
    std::deque<Req> lqueue;         // waiting for a lane (arrival order), owned by the host thread
    std::vector<kc_ctx*> lctx;      // one context per lane, open while the server is
    std::vector<std::thread> lanes;
    std::mutex lm;
    std::condition_variable lcv;
    std::deque<Req> ljobs;          // admitted, waiting for a worker
    struct LaneDone {
        kc_completion c;
        uint64_t base, charge;      // the arena charge to release
    };
    std::vector<LaneDone> ldone;
    uint32_t lbusy = 0;             // admitted large requests not yet reported (host thread)
    bool lquit = false;
};

static void lanes_stop(kc_server* s);

namespace {

std::mutex g_reg_m;  // open servers
std::vector<kc_server*> g_servers;
bool g_atexit = false;

struct HipFrees {
    static int dev_free(void* p) { return (int)hipFree(p); }
    static int host_free(void* p) { return (int)hipHostFree(p); }
    static int host_alloc(void** p, uint64_t bytes) { return (int)hipHostMalloc(p, bytes, hipHostMallocDefault); }
};
FreeRegistry<HipFrees> g_frees;  // resident serving kernels and the frees deferred meanwhile

kc_status free_status(int e) {
    return e == FreeRegistry<HipFrees>::kBadBlock ? KC_ERR_ARG : hip_status((hipError_t)e);
}

void resident_begin(kc_server* s) {
    g_frees.begin();
    s->resident = true;
}

// the server's kernel has ended: once none is resident, the deferred blocks are freed
void resident_end(kc_server* s) {
    if (!s->resident) return;
    s->resident = false;
    g_frees.end();
}

void hb_stop(kc_server* s) {
    if (!s->hb.joinable()) return;
    {
        std::lock_guard<std::mutex> l(s->hb_m);
        s->hb_quit = true;
    }
    s->hb_cv.notify_all();
    s->hb.join();
}

// Registered after the HIP runtime initialised, so it runs before the runtime's own teardown:
// stops every server still open when the process exits.
void serve_atexit() {
    std::vector<kc_server*> v;
    {
        std::lock_guard<std::mutex> l(g_reg_m);
        v = g_servers;
    }
    for (kc_server* s : v) {
        lanes_stop(s);
        if (!s->running) continue;
        __atomic_store_n(s->h_ctl + kHStop, 1u, __ATOMIC_RELEASE);
        hb_stop(s);
        const double t0 = now_s();
        while (hipStreamQuery(s->kstream) == hipErrorNotReady && now_s() - t0 < 5.0) std::this_thread::yield();
    }
}

void reg_add(kc_server* s) {
    std::lock_guard<std::mutex> l(g_reg_m);
    if (!g_atexit) {
        std::atexit(serve_atexit);
        g_atexit = true;
    }
    g_servers.push_back(s);
}

void reg_del(kc_server* s) {
    std::lock_guard<std::mutex> l(g_reg_m);
    g_servers.erase(std::remove(g_servers.begin(), g_servers.end(), s), g_servers.end());
}

}  // namespace

// A large-request lane: its own context (stream, events), one request at a time through the
// single-graph path with the whole state inside the request's arena region; results are
// reported through s->ldone.
static void lane_main(kc_server* s, kc_ctx* c) {
    (void)hipSetDevice(c->device);
    std::unique_lock<std::mutex> l(s->lm);
    while (true) {
        s->lcv.wait(l, [&] { return s->lquit || !s->ljobs.empty(); });
        if (s->ljobs.empty()) break;  // quit, and nothing left to run
        Req r = s->ljobs.front();
        s->ljobs.pop_front();
        l.unlock();
        kc_completion cp{};
        cp.req_id = r.id;
        cp.t_submit = r.t_submit;
        cp.t_admit = r.t_admit;
        kc_status st;
        {
            kc_csr h{r.V, r.arcs, r.off, r.nbr, nullptr};
            kc_graph* g = nullptr;
            st = graph_upload_into(c, &h, s->arena + r.base, r.charge, &g);
            if (st == KC_OK) {
                kc_params p{};
                p.flags = s->cfg.flags;
                kc_stats stt{};
                uint32_t kmax = 0;
                st = kc_peel(c, g, &p, r.core_out, &kmax, &stt);
                cp.kmax = kmax;
                cp.levels = stt.levels;
                cp.rounds = stt.subrounds;
                kc_graph_free(g);
            }
        }
        cp.status = st;
        cp.t_done = now_s();
        l.lock();
        s->ldone.push_back({cp, r.base, r.charge});
    }
}

static void lanes_stop(kc_server* s) {
    {
        std::lock_guard<std::mutex> l(s->lm);
        s->lquit = true;
    }
    s->lcv.notify_all();
    for (auto& t : s->lanes)
        if (t.joinable()) t.join();
    s->lanes.clear();
}

// charge of a large request: its whole single-graph state (offsets, neighbours, vertex arrays,
// chunk slots, alive lists, shell offsets, control and trace blocks), which the lane places in
// the request's arena region, and never less than the admission estimate 8 m + 16 n
static uint64_t large_charge(uint32_t V, uint32_t arcs, uint32_t dmax) {
    return std::max<uint64_t>(graph_bytes(V, arcs, dmax), 4ull * arcs + 16ull * V);
}

// Stops the serving kernel (the stop flag travels on the copy stream, after every publication,
// so every admitted request still runs) and frees everything. Returns KC_ERR_DEVICE when the
// kernel did not exit within the wait limit; its memory is then left allocated (the immediate
// stop flag is raised and the heartbeat stopped, so the kernel still ends).
static kc_status serve_release(kc_server* s) {
    lanes_stop(s);  // every admitted large request finishes first
    if (s->running) {
        *s->h_one = 1;
        hipError_t e = hipMemcpyAsync(s->d_head + kStopWord, s->h_one, 4, hipMemcpyHostToDevice, s->cstream);
        if (e != hipSuccess) __atomic_store_n(s->h_ctl + kHStop, 1u, __ATOMIC_RELEASE);
        const double t0 = now_s();
        while (hipStreamQuery(s->kstream) == hipErrorNotReady) {
            if (now_s() - t0 > s->ctx->wait_limit_ms * 1e-3) {
                __atomic_store_n(s->h_ctl + kHStop, 1u, __ATOMIC_RELEASE);
                hb_stop(s);
                return KC_ERR_DEVICE;
            }
            std::this_thread::yield();
        }
        s->running = false;
    }
    resident_end(s);
    hb_stop(s);
    reg_del(s);
    for (kc_ctx* c : s->lctx) kc_close(c);
    if (s->cstream) (void)hipStreamSynchronize(s->cstream);
    for (Req& r : s->slot)
        if (r.ev) (void)hipEventDestroy(r.ev);
    for (void* p : {(void*)s->arena, (void*)s->d_ring, (void*)s->d_head}) lib_free_device(p);
    for (void* p : {(void*)s->h_done, (void*)s->h_entry, (void*)s->h_ctl, (void*)s->h_one}) lib_free_pinned(p);
    if (s->kstream) (void)hipStreamDestroy(s->kstream);
    if (s->cstream) (void)hipStreamDestroy(s->cstream);
    delete s;
    return KC_OK;
}

// admit queued requests in arrival order while their slot is free and the arena has room
static kc_status serve_admit(kc_server* s) {
    // large requests: a free lane and room in the arena
    bool handed = false;
    while (!s->lqueue.empty() && s->lbusy < s->cfg.large_lanes) {
        Req& q = s->lqueue.front();
        uint64_t base;
        if (!s->alloc.alloc(q.charge, &base)) break;
        q.base = base;
        q.t_admit = now_s();
        {
            std::lock_guard<std::mutex> l(s->lm);
            s->ljobs.push_back(q);
        }
        s->lqueue.pop_front();
        s->lbusy++;
        handed = true;
    }
    if (handed) s->lcv.notify_all();
    const uint32_t R = s->cfg.ring_slots;
    s->fresh.clear();
    while (!s->queue.empty()) {
        // ring slots are reused in index order: the next index waits for its slot's last occupant
        const uint32_t sl = (uint32_t)(s->next_index % R);
        Req& r = s->slot[sl];
        if (r.state != 0) break;
        Req& q = s->queue.front();
        uint64_t base;
        if (!s->alloc.alloc(q.charge, &base)) break;
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
        KC_TRY(hipMemcpyAsync(dst + r.r.off, r.off, ((size_t)r.V + 1) * 4, hipMemcpyHostToDevice, s->cstream));
        if (r.arcs)
            KC_TRY(hipMemcpyAsync(dst + r.r.nbr, r.nbr, (size_t)r.arcs * 4, hipMemcpyHostToDevice, s->cstream));
        ServeEntry* e = s->h_entry + sl;
        e->base = base;
        e->V = r.V;
        e->arcs = r.arcs;
        e->flags = s->cfg.flags;
        e->pad = 0;
        e->out = r.out_pinned ? r.core_out : nullptr;
        e->pad2 = 0;
        e->seq = r.index + 1;
        s->fresh.push_back(sl);
    }
    // Entries go after all data on the same stream, so the kernel sees an entry only after its
    // request's data; consecutive slots are published by one copy.
    for (size_t i = 0; i < s->fresh.size();) {
        size_t j = i + 1;
        while (j < s->fresh.size() && s->fresh[j] == s->fresh[j - 1] + 1) ++j;
        KC_TRY(hipMemcpyAsync(s->d_ring + s->fresh[i], s->h_entry + s->fresh[i], (j - i) * sizeof(ServeEntry),
                              hipMemcpyHostToDevice, s->cstream));
        i = j;
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
    if (!s->cfg.lease_ms) s->cfg.lease_ms = 5000;
    if (!s->cfg.large_arcs) s->cfg.large_arcs = 1u << 20;
    if (!s->cfg.large_lanes) s->cfg.large_lanes = 2;
    if (s->cfg.block != 256 && s->cfg.block != 512 && s->cfg.block != 1024) {
        delete s;
        return KC_ERR_ARG;
    }
    // leave room on every CU for the runtime's own copy kernels next to the resident server
    if (!s->cfg.workgroups) s->cfg.workgroups = (uint32_t)c->cus * (s->cfg.block == 256 ? 2 : 1);
    const uint32_t R = s->cfg.ring_slots;
    reg_add(s);
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
    // completion records and the control page are shared live with the kernel: coherent memory
    if ((e = hipHostMalloc(&s->h_done, (size_t)R * sizeof(ServeDone), hipHostMallocCoherent))) return fail(e);
    if ((e = hipHostMalloc(&s->h_ctl, 256, hipHostMallocCoherent))) return fail(e);
    if ((e = hipHostMalloc(&s->h_entry, (size_t)R * sizeof(ServeEntry), hipHostMallocDefault))) return fail(e);
    if ((e = hipHostMalloc(&s->h_one, 256, hipHostMallocDefault))) return fail(e);
    memset(s->h_done, 0, (size_t)R * sizeof(ServeDone));
    memset(s->h_entry, 0, (size_t)R * sizeof(ServeEntry));
    memset(s->h_ctl, 0, 256);
    *s->h_one = 0;
    if ((e = hipMemsetAsync(s->d_ring, 0, (size_t)R * sizeof(ServeEntry), s->kstream))) return fail(e);
    if ((e = hipMemsetAsync(s->d_head, 0, 256, s->kstream))) return fail(e);
    // the ring must be clear before the copy stream can publish into it
    if ((e = hipStreamSynchronize(s->kstream))) return fail(e);
    s->slot.resize(R);
    for (Req& r : s->slot)
        if ((e = hipEventCreateWithFlags(&r.ev, hipEventDisableTiming))) return fail(e);
    s->alloc.init(s->cfg.hbm_budget_bytes);
    // heartbeat: advances at least 20 times per lease
    const uint32_t period_ms = std::max<uint32_t>(1, std::min<uint32_t>(50, s->cfg.lease_ms / 20));
    s->hb = std::thread([s, period_ms] {
        std::unique_lock<std::mutex> l(s->hb_m);
        while (!s->hb_quit) {
            __atomic_fetch_add(s->h_ctl + kHBeat, 1u, __ATOMIC_RELEASE);
            s->hb_cv.wait_for(l, std::chrono::milliseconds(period_ms));
        }
    });
    ServeArgs a{};
    a.arena = s->arena;
    a.ring = s->d_ring;
    a.done = s->h_done;
    a.head = reinterpret_cast<uint64_t*>(s->d_head);
    a.stop = s->d_head + kStopWord;
    a.hstop = s->h_ctl + kHStop;
    a.hbeat = s->h_ctl + kHBeat;
    a.expired = s->d_head + kExpired;
    a.lease_ticks = (uint64_t)s->cfg.lease_ms * c->wall_rate_khz;
    a.slots = R;
    if (s->cfg.large_arcs != 0xFFFFFFFFu)
        for (uint32_t i = 0; i < s->cfg.large_lanes; ++i) {
            kc_ctx* lc = nullptr;
            const kc_status os = kc_open(c->device, &lc);
            if (os != KC_OK) {
                (void)serve_release(s);
                return os;
            }
            s->lctx.push_back(lc);
        }
    KC_TRY(hipSetDevice(c->device));
    s->running = true;  // from here on the release path stops the kernel
    resident_begin(s);
    switch (s->cfg.block) {
        case 256: e = launch_serve<256>(a, s->cfg.workgroups, s->kstream); break;
        case 512: e = launch_serve<512>(a, s->cfg.workgroups, s->kstream); break;
        default: e = launch_serve<1024>(a, s->cfg.workgroups, s->kstream); break;
    }
    if (e) {
        s->running = false;
        resident_end(s);
        return fail(e);
    }
    for (kc_ctx* lc : s->lctx) s->lanes.emplace_back(lane_main, s, lc);
    *out = s;
    return KC_OK;
}

kc_status kc_serve_submit(kc_server* s, uint64_t req_id, const kc_csr* g, uint32_t* core_out) {
    if (!s || !g || !g->offsets || (g->arcs && !g->neighbors)) return KC_ERR_ARG;
    if (g->V == 0 || g->V >= (1u << 30) || g->arcs >= 0xFFFFFFFFull) return KC_ERR_CAPACITY;
    if (g->offsets[0] != 0 || g->offsets[g->V] != g->arcs) return KC_ERR_ARG;
    if (!s->running) return KC_ERR_DEVICE;
    Req q;
    q.id = req_id;
    q.V = g->V;
    q.arcs = (uint32_t)g->arcs;
    q.off = g->offsets;
    q.nbr = g->neighbors;
    q.core_out = core_out;
    q.out_pinned = core_out && is_pinned(core_out);
    q.large = !s->lanes.empty() && q.arcs >= s->cfg.large_arcs;
    q.r = serve_region(q.V, q.arcs);
    if (q.large) {
        uint32_t dmax = 0;
        for (uint32_t v = 0; v < q.V; ++v) {
            if (g->offsets[v + 1] < g->offsets[v]) return KC_ERR_ARG;
            dmax = std::max(dmax, g->offsets[v + 1] - g->offsets[v]);
        }
        q.charge = large_charge(q.V, q.arcs, dmax);
    } else {
        q.charge = std::max<uint64_t>(q.r.bytes, 4ull * q.arcs + 16ull * q.V);  // 8 m + 16 n
    }
    // every charge a multiple of 256 bytes: with the arena's first fit from offset 0, every
    // region then starts on a 256-byte boundary (graph_upload_into requires it)
    q.charge = serve_align(q.charge);
    if (q.charge > s->cfg.hbm_budget_bytes) return KC_ERR_CAPACITY;
    q.t_submit = now_s();
    (q.large ? s->lqueue : s->queue).push_back(q);
    return serve_admit(s);
}

kc_status kc_serve_poll(kc_server* s, kc_completion* out, uint32_t cap, uint32_t* n_out) {
    if (!s || !n_out || (cap && !out)) return KC_ERR_ARG;
    *n_out = 0;
    if (!s->running) return KC_ERR_DEVICE;
    uint32_t n = 0;
    if (s->lbusy) {
        std::lock_guard<std::mutex> l(s->lm);
        while (!s->ldone.empty() && n < cap) {
            const kc_server::LaneDone d = s->ldone.back();
            s->ldone.pop_back();
            out[n++] = d.c;
            s->alloc.release(d.base, d.charge);
            s->lbusy--;
        }
    }
    for (uint32_t sl = 0; sl < s->cfg.ring_slots && s->inflight; ++sl) {
        Req& r = s->slot[sl];
        if (r.state == 1) {
            const uint64_t seq = __atomic_load_n(&s->h_done[sl].seq, __ATOMIC_ACQUIRE);
            if (seq != r.index + 1) continue;
            if (r.core_out && !r.out_pinned) {
                KC_TRY(hipMemcpyAsync(r.core_out, s->arena + r.base + r.r.core, (size_t)r.V * 4,
                                      hipMemcpyDeviceToHost, s->cstream));
                KC_TRY(hipEventRecord(r.ev, s->cstream));
            }
            r.state = 2;
        }
        if (r.state == 2 && n < cap) {
            if (r.core_out && !r.out_pinned) {
                const hipError_t q = hipEventQuery(r.ev);
                if (q == hipErrorNotReady) continue;
                KC_TRY(q);
            }
            kc_completion& c = out[n++];
            const ServeDone& d = s->h_done[sl];
            c.req_id = r.id;
            c.kmax = d.kmax;
            c.levels = d.levels;
            c.rounds = d.rounds;
            c.status = d.err == SERVE_OK ? KC_OK : (d.err == SERVE_BAD_CSR ? KC_ERR_ARG : KC_ERR_NO_PROGRESS);
            c.t_submit = r.t_submit;
            c.t_admit = r.t_admit;
            c.t_done = now_s();
            s->alloc.release(r.base, r.charge);
            r.state = 0;
            s->inflight--;
        }
    }
    *n_out = n;
    // a kernel that ended while the server is open (its lease ran out) is reported, not waited on
    if (n == 0 && s->inflight && hipStreamQuery(s->kstream) == hipSuccess) {
        s->running = false;
        resident_end(s);
        return KC_ERR_DEVICE;
    }
    return serve_admit(s);
}

kc_status kc_serve_pending(const kc_server* s, uint32_t* queued, uint32_t* inflight, uint64_t* arena_used) {
    if (!s) return KC_ERR_ARG;
    if (queued) *queued = (uint32_t)(s->queue.size() + s->lqueue.size());
    if (inflight) *inflight = s->inflight + s->lbusy;
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
    return free_status(g_frees.alloc_pinned(bytes, out));
}

kc_status kc_host_free(void* p) { return free_status(g_frees.free_pinned(p)); }

}  // extern "C"

void lib_free_device(void* p) { (void)g_frees.free_device(p); }

void lib_free_pinned(void* p) { (void)g_frees.free_host(p); }
