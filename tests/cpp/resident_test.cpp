// Checks of the deferred-free registry (kcoregpu_amd/csrc/resident.h) with a malloc-backed fake
// backend: no raw free ever runs while a serving kernel counts as resident; frees made meanwhile
// are kept and all released when the last kernel ends; returned pinned blocks are handed out again
// only when at least as large as asked and at most twice it; foreign pointers and second frees are
// refused; nothing leaks. Then the same invariants under several threads at once (built with ASan
// or TSan by the Python test).
#include <atomic>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <set>
#include <thread>
#include <vector>

#include "../../kcoregpu_amd/csrc/resident.h"

static std::atomic<int> g_running{0};     // kernels between begin() returning and end() being called
static std::atomic<long> g_violations{0};  // raw frees while a kernel ran
static std::atomic<long> g_dev_frees{0}, g_host_frees{0}, g_host_allocs{0};
static std::atomic<long> g_foreign{0};     // free_host calls with blocks not from alloc_pinned

struct Fake {
    static int dev_free(void* p) {
        if (g_running.load()) ++g_violations;
        ++g_dev_frees;
        std::free(p);
        return 0;
    }
    static int host_free(void* p) {
        if (g_running.load()) ++g_violations;
        ++g_host_frees;
        std::free(p);
        return 0;
    }
    static int host_alloc(void** p, uint64_t bytes) {
        *p = std::malloc(bytes);
        if (!*p) return 2;
        ++g_host_allocs;
        return 0;
    }
};

using Reg = kcdev::FreeRegistry<Fake>;

#define CHECK(c)                                                             \
    do {                                                                     \
        if (!(c)) {                                                          \
            std::printf("FAIL %s:%d: %s\n", __FILE__, __LINE__, #c);         \
            std::exit(1);                                                    \
        }                                                                    \
    } while (0)

static void* pinned(Reg& r, uint64_t n) {
    void* p = nullptr;
    CHECK(r.alloc_pinned(n, &p) == 0 && p);
    std::memset(p, 0x5A, n);  // the whole size asked for is usable
    return p;
}

static void single_thread() {
    Reg r;
    // nothing resident: frees are immediate
    CHECK(r.free_device(std::malloc(64)) == 0 && g_dev_frees == 1);
    void* a = pinned(r, 5000);
    CHECK(r.free_pinned(a) == 0 && g_host_frees == 1);
    CHECK(r.free_pinned(a) == Reg::kBadBlock);  // freed for good: no longer a block of the registry
    int local = 0;
    CHECK(r.free_pinned(&local) == Reg::kBadBlock);

    r.begin();
    ++g_running;
    void* d = std::malloc(128);
    CHECK(r.free_device(d) == 0 && g_dev_frees == 1);
    void* h = std::malloc(256);  // pinned, but not from alloc_pinned
    ++g_foreign;
    CHECK(r.free_host(h) == 0 && g_host_frees == 1);
    void* b = pinned(r, 5000);   // 8 KiB block
    CHECK(r.free_pinned(b) == 0);
    CHECK(r.free_pinned(b) == Reg::kBadBlock);  // second free of a kept block
    CHECK(pinned(r, 6000) == b);                // handed out again
    CHECK(r.free_pinned(b) == 0);
    CHECK(pinned(r, 4096) == b);                // 8 KiB <= 2 x 4 KiB
    CHECK(r.free_pinned(b) == 0);
    void* c = pinned(r, 3 * 4096);              // nothing kept is large enough: new block
    CHECK(c != b);
    void* e = pinned(r, 100);                   // 4 KiB asked, 8 KiB kept: reused
    CHECK(e == b);
    CHECK(r.free_pinned(c) == 0);               // 12 KiB kept
    void* f = pinned(r, 4096);                  // 12 KiB > 2 x 4 KiB: new block
    CHECK(f != c && f != b);
    CHECK(r.free_pinned(e) == 0 && r.free_pinned(f) == 0);
    CHECK(r.free_host(c) == Reg::kBadBlock);    // free_host of a kept alloc_pinned block
    CHECK(r.kept() == 5);                       // d, h, and the blocks b, c, f
    CHECK(g_dev_frees == 1 && g_host_frees == 1);

    // a second kernel: the first one ending frees nothing
    r.begin();
    ++g_running;
    --g_running;
    r.end();
    CHECK(r.kept() == 5 && r.resident() == 1);
    --g_running;
    r.end();
    CHECK(r.kept() == 0 && r.resident() == 0);
    CHECK(g_dev_frees == 2 && g_host_frees == 5);
    r.end();  // unmatched: ignored
    CHECK(r.resident() == 0);
    CHECK(r.free_pinned(b) == Reg::kBadBlock);  // released at the end: gone

    // blocks live across the end are still the registry's
    void* g = pinned(r, 1 << 20);
    r.begin();
    ++g_running;
    --g_running;
    r.end();
    CHECK(r.free_pinned(g) == 0 && g_host_frees == 6);
    CHECK(g_violations == 0);
}

static void threads(int nthreads, int iters) {
    Reg r;
    std::vector<std::thread> th;
    std::atomic<long> ops{0};
    for (int t = 0; t < nthreads; ++t)
        th.emplace_back([&, t] {
            std::mt19937 rng(1234 + t);
            std::vector<std::pair<void*, uint64_t>> mine;
            int sessions = 0;
            for (int i = 0; i < iters; ++i) {
                const int op = (int)(rng() % 8);
                if (op <= 2) {
                    const uint64_t n = 1 + rng() % 40000;
                    void* p = pinned(r, n);
                    mine.emplace_back(p, n);
                } else if (op <= 4 && !mine.empty()) {
                    const size_t k = rng() % mine.size();
                    std::memset(mine[k].first, 0x33, mine[k].second);
                    CHECK(r.free_pinned(mine[k].first) == 0);
                    mine[k] = mine.back();
                    mine.pop_back();
                } else if (op == 5) {
                    CHECK(r.free_device(std::malloc(1 + rng() % 512)) == 0);
                } else if (op == 6) {
                    ++g_foreign;
                    CHECK(r.free_host(std::malloc(1 + rng() % 512)) == 0);
                } else if (sessions < iters / 16) {
                    ++sessions;
                    r.begin();
                    ++g_running;
                    std::this_thread::yield();
                    --g_running;
                    r.end();
                }
                ++ops;
            }
            for (auto& m : mine) CHECK(r.free_pinned(m.first) == 0);
        });
    for (auto& t : th) t.join();
    CHECK(r.resident() == 0);
    CHECK(r.kept() == 0);  // the last end() released everything
    CHECK(g_violations == 0);
    std::printf("threads %d ops %ld\n", nthreads, ops.load());
}

int main(int argc, char** argv) {
    const int iters = argc > 1 ? std::atoi(argv[1]) : 20000;
    single_thread();
    threads(8, iters);
    CHECK(g_host_allocs + g_foreign == g_host_frees);
    std::printf("ok host_allocs %ld\n", g_host_allocs.load());
    return 0;
}
