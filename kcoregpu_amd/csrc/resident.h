// resident.h — deferred frees while a serving kernel is resident.
//
// hipFree and hipHostFree synchronise the whole device: called while a resident serving kernel
// runs, they wait for it, i.e. until the server is closed. FreeRegistry counts the resident
// kernels and keeps every free the library makes while one runs; when the last one ends, the
// kept blocks are freed. Pinned host blocks handed out by alloc_pinned and returned meanwhile are
// handed out again (the smallest kept block of at least the size asked for and at most twice it),
// so a serving loop that allocates one result buffer per request reaches a steady state.
//
// The backend B supplies the raw calls, each returning 0 on success or an error code:
//   static int dev_free(void*); static int host_free(void*); static int host_alloc(void**, uint64_t);
// One mutex guards the registry and is held across every raw call, and begin() takes it before a
// launch, so a free never starts while a kernel that it would wait for is resident. Thread-safe.
#pragma once
#include <cstdint>
#include <map>
#include <mutex>
#include <unordered_map>
#include <vector>

namespace kcdev {

template <class B>
class FreeRegistry {
public:
    static constexpr int kBadBlock = -1;  // free_pinned: not an alloc_pinned block, or freed twice

    // A serving kernel is about to be launched.
    void begin() {
        std::lock_guard<std::mutex> l(m_);
        ++resident_;
    }
    // A serving kernel counted by begin() has ended (once per begin).
    void end() {
        std::lock_guard<std::mutex> l(m_);
        if (resident_ == 0 || --resident_) return;
        for (void* p : dev_later_) B::dev_free(p);
        for (void* p : host_later_) B::host_free(p);
        dev_later_.clear();
        host_later_.clear();
        for (auto& kv : idle_) {
            B::host_free(kv.second);
            pin_.erase(kv.second);
        }
        idle_.clear();
    }
    // Device memory from hipMalloc.
    int free_device(void* p) {
        if (!p) return 0;
        std::lock_guard<std::mutex> l(m_);
        if (resident_) {
            dev_later_.push_back(p);
            return 0;
        }
        return B::dev_free(p);
    }
    // Pinned host memory: an alloc_pinned block, or any other hipHostMalloc block.
    int free_host(void* p) {
        if (!p) return 0;
        std::lock_guard<std::mutex> l(m_);
        auto it = pin_.find(p);
        if (it != pin_.end()) return release(it);
        if (resident_) {
            host_later_.push_back(p);
            return 0;
        }
        return B::host_free(p);
    }
    // Pinned host block of at least `bytes` (rounded up to 4 KiB).
    int alloc_pinned(uint64_t bytes, void** out) {
        *out = nullptr;
        if (!bytes) return 0;
        const uint64_t sz = (bytes + 4095) & ~4095ull;
        std::lock_guard<std::mutex> l(m_);
        auto it = idle_.lower_bound(sz);
        if (it != idle_.end() && it->first <= 2 * sz) {
            *out = it->second;
            pin_[it->second].idle = false;
            idle_.erase(it);
            return 0;
        }
        void* p = nullptr;
        const int e = B::host_alloc(&p, sz);
        if (e) return e;
        pin_[p] = Block{sz, false};
        *out = p;
        return 0;
    }
    // Returns an alloc_pinned block; kBadBlock for any other pointer or a second free.
    int free_pinned(void* p) {
        if (!p) return 0;
        std::lock_guard<std::mutex> l(m_);
        auto it = pin_.find(p);
        if (it == pin_.end()) return kBadBlock;
        return release(it);
    }

    // for tests
    uint32_t resident() {
        std::lock_guard<std::mutex> l(m_);
        return resident_;
    }
    size_t kept() {
        std::lock_guard<std::mutex> l(m_);
        return dev_later_.size() + host_later_.size() + idle_.size();
    }

private:
    struct Block {
        uint64_t bytes;
        bool idle;
    };
    int release(typename std::unordered_map<void*, Block>::iterator it) {
        if (it->second.idle) return kBadBlock;
        if (resident_) {
            it->second.idle = true;
            idle_.emplace(it->second.bytes, it->first);
            return 0;
        }
        void* p = it->first;
        pin_.erase(it);
        return B::host_free(p);
    }

    std::mutex m_;
    uint32_t resident_ = 0;
    std::vector<void*> dev_later_, host_later_;
    std::unordered_map<void*, Block> pin_;     // live alloc_pinned blocks
    std::multimap<uint64_t, void*> idle_;      // returned while a kernel was resident, by size
};

}  // namespace kcdev
