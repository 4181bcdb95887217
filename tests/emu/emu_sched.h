// emu_sched.h — the wave64 emulator's scheduler (test infrastructure; see hip/hip_runtime.h).
//
// launch(grid, block, fn) runs `grid` workgroups concurrently, one OS thread each; inside a
// workgroup every work-item is a ucontext fiber. A fiber runs until it reaches a wave-wide
// operation, a workgroup barrier or a sleep, then hands control back to its workgroup's scheduler:
//  * wave operations complete once every lane of the wave that has not exited waits at one; all
//    of them must be the same operation (a wave operation reached by only some lanes of a wave is
//    reported as an error, since this project's kernels never rely on partial-wave semantics);
//  * a workgroup barrier completes once every work-item that has not exited waits at it;
//  * a sleep just yields. When a whole pass finds nothing but sleepers, the OS thread yields.
// In weak-memory mode (KCEMU_WEAK=1) each work-item has a store buffer (see hip/hip_runtime.h):
// at every yield each buffered store completes with probability 1/3 (in random order); a sleep,
// the end of the work-item and a pass in which nothing could run complete them all.
// KCEMU_JITTER=N[:US]: after a scheduling pass, a workgroup pauses for up to US microseconds
// (default 200) with probability 1/N, so workgroups overtake each other at random points (more
// cross-workgroup interleavings); pauses of milliseconds stand for a preempted workgroup.
// Deadlocks inside a workgroup (every live fiber blocked on something that cannot complete)
// abort with a message; waits on other workgroups are the kernels' own, bounded by their wall
// clock limits (clock_ticks() counts 10 ns ticks).
#pragma once
#include <ucontext.h>

#include <atomic>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <thread>
#include <vector>

#include "hip/hip_runtime.h"

namespace kcemu {

enum State { RUN = 0, WAIT_WAVE, WAIT_BAR, DONE };

struct SbEntry {
    void* p;
    uint64_t v;
    int size;
};

struct Fiber {
    ucontext_t ctx;
    std::vector<char> stack;
    int state = RUN;
    int op = 0;
    uint64_t a = 0, b = 0, r = 0;
    std::vector<SbEntry> sb;  // weak-memory store buffer, program order
    uint64_t rng = 0;
};

struct Group {
    uint32_t bid = 0, grid = 0, block = 0;
    std::vector<Fiber> f;
    ucontext_t sched;
    int cur = -1;
    int bar_or = 0;
    std::function<void()> body;
};

inline thread_local Group* g_grp = nullptr;
inline const bool g_weak = [] {
    const char* e = std::getenv("KCEMU_WEAK");
    return e && *e && *e != '0';
}();
inline const uint32_t g_jitter = [] {
    const char* e = std::getenv("KCEMU_JITTER");
    return e && *e ? (uint32_t)std::strtoul(e, nullptr, 10) : 0u;
}();
inline const uint32_t g_jitter_us = [] {
    const char* e = std::getenv("KCEMU_JITTER");
    const char* c = e ? std::strchr(e, ':') : nullptr;
    const uint32_t us = c ? (uint32_t)std::strtoul(c + 1, nullptr, 10) : 0u;
    return us ? us : 200u;
}();

inline void sb_complete(const SbEntry& e) {
    switch (e.size) {
        case 1: __atomic_store_n(static_cast<uint8_t*>(e.p), (uint8_t)e.v, __ATOMIC_SEQ_CST); break;
        case 2: __atomic_store_n(static_cast<uint16_t*>(e.p), (uint16_t)e.v, __ATOMIC_SEQ_CST); break;
        case 4: __atomic_store_n(static_cast<uint32_t*>(e.p), (uint32_t)e.v, __ATOMIC_SEQ_CST); break;
        default: __atomic_store_n(static_cast<uint64_t*>(e.p), e.v, __ATOMIC_SEQ_CST); break;
    }
}
inline void sb_drain_fiber(Fiber& f) {
    for (const SbEntry& e : f.sb) sb_complete(e);
    f.sb.clear();
}
inline uint64_t sb_rand(Fiber& f) {
    f.rng ^= f.rng << 13;
    f.rng ^= f.rng >> 7;
    f.rng ^= f.rng << 17;
    return f.rng;
}
// some buffered stores complete: each with probability 1/3, the rest keep their order
inline void sb_some(Fiber& f) {
    size_t k = 0;
    for (size_t i = 0; i < f.sb.size(); ++i) {
        if (sb_rand(f) % 3 == 0)
            sb_complete(f.sb[i]);
        else
            f.sb[k++] = f.sb[i];
    }
    f.sb.resize(k);
}
inline Fiber* sb_self() { return (g_weak && g_grp && g_grp->cur >= 0) ? &g_grp->f[g_grp->cur] : nullptr; }
inline bool overlaps(const SbEntry& e, const void* p, int size) {
    const char* a = static_cast<const char*>(e.p);
    const char* b = static_cast<const char*>(p);
    return a < b + size && b < a + e.size;
}

inline bool sb_put(void* p, uint64_t v, int size) {
    Fiber* f = sb_self();
    if (!f) return false;
    for (SbEntry& e : f->sb)
        if (overlaps(e, p, size)) {
            if (e.p == p && e.size == size) {  // a later store to the same word replaces it
                e.v = v;
                return true;
            }
            sb_drain_fiber(*f);
            break;
        }
    f->sb.push_back(SbEntry{p, v, size});
    return true;
}
inline bool sb_get(const void* p, int size, uint64_t* v) {
    Fiber* f = sb_self();
    if (!f) return false;
    for (const SbEntry& e : f->sb)
        if (overlaps(e, p, size)) {
            if (e.p == p && e.size == size) {
                *v = e.v;
                return true;
            }
            sb_drain_fiber(*f);
            return false;
        }
    return false;
}
inline void sb_touch(const void* p, int size) {
    Fiber* f = sb_self();
    if (!f) return;
    for (const SbEntry& e : f->sb)
        if (overlaps(e, p, size)) {
            sb_drain_fiber(*f);
            return;
        }
}
inline void sb_drain() {
    if (Fiber* f = sb_self()) sb_drain_fiber(*f);
}

inline kcemu_dim3 tid3() { return kcemu_dim3{(uint32_t)g_grp->cur, 0, 0}; }
inline kcemu_dim3 bid3() { return kcemu_dim3{g_grp->bid, 0, 0}; }
inline kcemu_dim3 bdim3() { return kcemu_dim3{g_grp->block, 1, 1}; }
inline kcemu_dim3 gdim3() { return kcemu_dim3{g_grp->grid, 1, 1}; }
inline uint32_t lane() { return (uint32_t)g_grp->cur & 63u; }

inline void yield_to_sched() {
    Fiber& me = g_grp->f[g_grp->cur];
    if (g_weak) sb_some(me);
    swapcontext(&me.ctx, &g_grp->sched);
}

inline uint64_t coll(int op, uint64_t x, uint64_t y) {
    Fiber& me = g_grp->f[g_grp->cur];
    me.state = WAIT_WAVE;
    me.op = op;
    me.a = x;
    me.b = y;
    yield_to_sched();
    return me.r;
}

inline void sync_block() {
    Fiber& me = g_grp->f[g_grp->cur];
    me.state = WAIT_BAR;
    me.a = 0;
    yield_to_sched();
}

inline int sync_block_or(int x) {
    Fiber& me = g_grp->f[g_grp->cur];
    me.state = WAIT_BAR;
    me.a = x ? 1 : 0;
    yield_to_sched();
    return (int)me.r;
}

inline void nap() {
    sb_drain();  // a spinning work-item's stores complete eventually
    yield_to_sched();
}

inline uint64_t clock_ticks() {
    using namespace std::chrono;
    return (uint64_t)duration_cast<nanoseconds>(steady_clock::now().time_since_epoch()).count() / 10;
}

[[noreturn]] inline void fail(const char* what, uint32_t bid, int t) {
    std::fprintf(stderr, "emu: %s (workgroup %u, work-item %d)\n", what, bid, t);
    std::fflush(stderr);
    std::abort();
}

// resolve one wave whose live lanes all wait at a wave operation
inline void resolve_wave(Group& g, int w) {
    const int base = w * 64;
    const int n = (int)std::min<uint32_t>(64, g.block - base);
    int op = -1;
    uint64_t active = 0;
    uint32_t val[64] = {0};
    for (int l = 0; l < n; ++l) {
        Fiber& f = g.f[base + l];
        if (f.state == DONE) continue;
        if (op < 0) op = f.op;
        if (f.op != op) {
            std::fprintf(stderr, "emu: wave %d ops:", w);
            for (int k = 0; k < n; ++k) std::fprintf(stderr, " %d:%d/%d", k, g.f[base + k].state, g.f[base + k].op);
            std::fprintf(stderr, "\n");
            fail("lanes of one wave wait at different wave operations", g.bid, base + l);
        }
        active |= 1ull << l;
        val[l] = (uint32_t)f.a;
    }
    uint64_t ballot = 0;
    for (int l = 0; l < n; ++l)
        if ((active >> l & 1) && g.f[base + l].a) ballot |= 1ull << l;
    int first = -1;
    for (int l = 0; l < n && first < 0; ++l)
        if (active >> l & 1) first = l;
    for (int l = 0; l < n; ++l) {
        if (!(active >> l & 1)) continue;
        Fiber& f = g.f[base + l];
        uint64_t r = 0;
        switch (op) {
            case OP_BALLOT: r = ballot; break;
            case OP_FIRSTLANE: r = val[first]; break;
            case OP_READLANE: r = val[(int)f.b & 63]; break;
            case OP_SHFL_XOR: r = val[(l ^ (int)f.b) & 63]; break;
            case OP_WAVE_BARRIER: r = 0; break;
            case OP_DPP: {
                const uint32_t ctrl = (uint32_t)(f.b & 0xFFFF), row_mask = (uint32_t)((f.b >> 16) & 0xF);
                const uint32_t old = (uint32_t)(f.b >> 32);
                const int row = l / 16, in_row = l % 16;
                r = old;
                if (!((row_mask >> row) & 1)) break;
                if (ctrl >= 0x111 && ctrl <= 0x11F) {  // row_shr:n
                    const int s = (int)(ctrl - 0x110);
                    if (in_row >= s) r = val[l - s];
                } else if (ctrl == 0x142) {  // row_bcast:15 -> rows after row 0 get lane 15 of the previous row
                    if (row > 0) r = val[row * 16 - 1];
                } else if (ctrl == 0x143) {  // row_bcast:31 -> rows 2 and 3 get lane 31
                    if (row >= 2) r = val[31];
                } else {
                    fail("unsupported DPP control", g.bid, base + l);
                }
                break;
            }
            default: fail("unknown wave operation", g.bid, base + l);
        }
        f.r = r;
        f.state = RUN;
    }
}

inline void trampoline(uint32_t hi, uint32_t lo) {
    Group* g = reinterpret_cast<Group*>(((uintptr_t)hi << 32) | lo);
    g->body();
    sb_drain_fiber(g->f[g->cur]);  // a kernel's stores are complete when it ends
    g->f[g->cur].state = DONE;
    swapcontext(&g->f[g->cur].ctx, &g->sched);
}

inline void run_group(Group& g) {
    g_grp = &g;
    uint64_t jitter_state = 0x2545F4914F6CDD1Dull ^ ((uint64_t)g.bid * 0x9E3779B97F4A7C15ull) ^ (uint64_t)clock_ticks();
    const size_t stack = 128 * 1024;
    for (uint32_t t = 0; t < g.block; ++t) {
        Fiber& f = g.f[t];
        f.stack.resize(stack);
        getcontext(&f.ctx);
        f.ctx.uc_stack.ss_sp = f.stack.data();
        f.ctx.uc_stack.ss_size = stack;
        f.ctx.uc_link = nullptr;
        const uintptr_t p = reinterpret_cast<uintptr_t>(&g);
        makecontext(&f.ctx, (void (*)())trampoline, 2, (uint32_t)(p >> 32), (uint32_t)p);
        f.rng = 0x9E3779B97F4A7C15ull ^ ((uint64_t)g.bid << 32) ^ (uint64_t)(t + 1) * 0xD1B54A32D192ED03ull;
    }
    const int nw = (int)((g.block + 63) / 64);
    while (true) {
        bool any_live = false, progressed = false;
        for (uint32_t t = 0; t < g.block; ++t) {
            Fiber& f = g.f[t];
            if (f.state != RUN) continue;
            any_live = true;
            g.cur = (int)t;
            swapcontext(&g.sched, &f.ctx);
            if (f.state != RUN) progressed = true;
        }
        // wave operations
        for (int w = 0; w < nw; ++w) {
            bool all_wait = true, any = false;
            for (int l = 0; l < 64 && w * 64 + l < (int)g.block; ++l) {
                const int s = g.f[w * 64 + l].state;
                if (s == DONE) continue;
                any = true;
                if (s != WAIT_WAVE) all_wait = false;
            }
            if (any && all_wait) {
                resolve_wave(g, w);
                progressed = true;
            }
        }
        // workgroup barrier
        bool bar_all = true, bar_any = false;
        int acc = 0;
        for (uint32_t t = 0; t < g.block; ++t) {
            const int s = g.f[t].state;
            if (s == DONE) continue;
            any_live = true;
            if (s != WAIT_BAR) bar_all = false;
            else {
                bar_any = true;
                acc |= (int)g.f[t].a;
            }
        }
        if (bar_any && bar_all) {
            for (uint32_t t = 0; t < g.block; ++t)
                if (g.f[t].state == WAIT_BAR) {
                    g.f[t].state = RUN;
                    g.f[t].r = (uint64_t)acc;
                }
            progressed = true;
        }
        bool all_done = true;
        for (uint32_t t = 0; t < g.block; ++t)
            if (g.f[t].state != DONE) all_done = false;
        if (all_done) break;
        bool any_run = false;
        for (uint32_t t = 0; t < g.block; ++t)
            if (g.f[t].state == RUN) any_run = true;
        if (!any_run && !progressed) {
            // everyone waits: a barrier some work-items will never reach, or waves split
            for (uint32_t t = 0; t < g.block; ++t)
                if (g.f[t].state == WAIT_BAR) fail("workgroup barrier not reached by every live work-item", g.bid, (int)t);
            fail("wave operation never completes", g.bid, 0);
        }
        if (!progressed) {
            // nothing could run: every buffered store completes (the GPU's do, eventually)
            for (uint32_t t = 0; t < g.block; ++t) sb_drain_fiber(g.f[t]);
            std::this_thread::yield();
        }
        if (g_jitter) {
            uint64_t& r = jitter_state;
            r ^= r << 13;
            r ^= r >> 7;
            r ^= r << 17;
            if (r % g_jitter == 0) std::this_thread::sleep_for(std::chrono::microseconds((r >> 20) % g_jitter_us));
        }
        (void)any_live;
    }
    g_grp = nullptr;
}

// Run fn() as a kernel of `grid` workgroups of `block` work-items, every workgroup at once.
inline void launch(uint32_t grid, uint32_t block, const std::function<void()>& fn) {
    std::vector<Group> groups(grid);
    std::vector<std::thread> th;
    for (uint32_t b = 0; b < grid; ++b) {
        groups[b].bid = b;
        groups[b].grid = grid;
        groups[b].block = block;
        groups[b].f.resize(block);
        groups[b].body = fn;
    }
    for (uint32_t b = 0; b < grid; ++b) th.emplace_back([&, b] { run_group(groups[b]); });
    for (auto& t : th) t.join();
}

}  // namespace kcemu
