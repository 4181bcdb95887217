// arena.h — first-fit allocator of byte ranges inside one device allocation (the serving arena).
// Host-only bookkeeping: free ranges kept by offset, neighbours merged on release.
#pragma once
#include <cstdint>
#include <iterator>
#include <map>

namespace kcdev {

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

}  // namespace kcdev
