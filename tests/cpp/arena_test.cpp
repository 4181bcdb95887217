// Randomised check of the serving arena's first-fit allocator (kcoregpu_amd/csrc/arena.h) against
// a brute-force model: allocations stay inside the arena and never overlap, each one is placed at
// the lowest offset with a large enough gap, an allocation fails exactly when no gap is large
// enough, `used` is the sum of live sizes, and releasing everything leaves one free range.
#include <cstdio>
#include <cstdlib>
#include <map>
#include <random>
#include <vector>

#include "../../kcoregpu_amd/csrc/arena.h"

static uint64_t model_first_fit(const std::map<uint64_t, uint64_t>& live, uint64_t size, uint64_t n, bool* ok) {
    uint64_t pos = 0;
    for (const auto& kv : live) {
        if (kv.first - pos >= n) {
            *ok = true;
            return pos;
        }
        pos = kv.first + kv.second;
    }
    *ok = size - pos >= n;
    return pos;
}

int main(int argc, char** argv) {
    const int seeds = argc > 1 ? std::atoi(argv[1]) : 20;
    long checks = 0;
    for (int seed = 0; seed < seeds; ++seed) {
        std::mt19937_64 rng(seed);
        const uint64_t size = 1 + rng() % 100000;
        kcdev::Arena a;
        a.init(size);
        std::map<uint64_t, uint64_t> live;  // offset -> length
        for (int step = 0; step < 4000; ++step) {
            if (live.empty() || rng() % 3) {
                const uint64_t n = 1 + rng() % (size / (1 + rng() % 16) + 1);
                bool mok = false;
                const uint64_t want = model_first_fit(live, size, n, &mok);
                uint64_t off = ~0ull;
                const bool ok = a.alloc(n, &off);
                if (ok != mok || (ok && off != want)) {
                    std::printf("seed %d step %d: alloc(%llu) -> %d @%llu, model %d @%llu\n", seed, step,
                                (unsigned long long)n, ok, (unsigned long long)off, mok, (unsigned long long)want);
                    return 1;
                }
                if (ok) live[off] = n;
            } else {
                auto it = live.begin();
                std::advance(it, rng() % live.size());
                a.release(it->first, it->second);
                live.erase(it);
            }
            uint64_t sum = 0, end = 0;
            for (const auto& kv : live) {
                if (kv.first < end || kv.first + kv.second > size) {
                    std::printf("seed %d step %d: overlap or out of range\n", seed, step);
                    return 1;
                }
                end = kv.first + kv.second;
                sum += kv.second;
            }
            if (sum != a.used) {
                std::printf("seed %d step %d: used %llu, live %llu\n", seed, step, (unsigned long long)a.used,
                            (unsigned long long)sum);
                return 1;
            }
            ++checks;
        }
        for (const auto& kv : live) a.release(kv.first, kv.second);
        if (a.used != 0 || a.free_.size() != 1 || a.free_.begin()->first != 0 || a.free_.begin()->second != size) {
            std::printf("seed %d: free list not merged back to one range\n", seed);
            return 1;
        }
    }
    std::printf("ok %ld checks\n", checks);
    return 0;
}
