// Device memory for the library's buffers (DevBuf): a per-device cache of
// released blocks, so that a search or batch that allocates and releases
// the same sizes call after call maps memory once.
//
// A release keeps hipFree's synchronisation -- the device finishes all its
// work before the block is cached -- so a cached block is never handed out
// while a kernel or copy enqueued before its release may still use it, and
// code after a release sees completed work exactly as after hipFree. What is
// saved is the unmapping and mapping. Sizes are rounded (powers of two up to
// 1 MiB, then multiples of 2 MiB); a request takes the smallest cached block
// of at least its size and at most 5/4 of it. At most kCacheCap bytes stay
// cached per device; beyond that a release frees the block.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <map>
#include <mutex>
#include <unordered_map>

#include "ctx.h"

namespace saber {

namespace {

constexpr uint64_t kCacheCap = 32ull << 30;
constexpr int kMaxDevices = 64;

struct DevCache {
    std::mutex mu;
    std::multimap<uint64_t, void*> free_blocks[kMaxDevices];   // size -> block
    uint64_t cached[kMaxDevices] = {};
    struct Live {
        int device;
        uint64_t size;
    };
    std::unordered_map<void*, Live> live;
};

DevCache& cache() {
    static DevCache* c = new DevCache();   // lives until the process ends
    return *c;
}

uint64_t rounded(uint64_t bytes) {
    if (bytes <= (1ull << 20)) {
        uint64_t s = 256;
        while (s < bytes) s <<= 1;
        return s;
    }
    const uint64_t g = 2ull << 20;
    return (bytes + g - 1) / g * g;
}

}  // namespace

hipError_t dev_alloc(void** p, uint64_t bytes) {
    *p = nullptr;
    int dev = 0;
    if (hipError_t e = hipGetDevice(&dev); e != hipSuccess) return e;
    const uint64_t size = rounded(bytes);
    DevCache& C = cache();
    if (dev >= 0 && dev < kMaxDevices) {
        std::lock_guard<std::mutex> g(C.mu);
        auto& F = C.free_blocks[dev];
        auto it = F.lower_bound(size);
        if (it != F.end() && it->first <= size + size / 4) {
            *p = it->second;
            C.cached[dev] -= it->first;
            C.live[*p] = DevCache::Live{dev, it->first};
            F.erase(it);
            return hipSuccess;
        }
    }
    hipError_t e = hipMalloc(p, size);
    if (e == hipErrorOutOfMemory && dev >= 0 && dev < kMaxDevices) {
        // give the cached blocks back and try once more
        (void)hipGetLastError();
        std::multimap<uint64_t, void*> drop;
        {
            std::lock_guard<std::mutex> g(C.mu);
            drop.swap(C.free_blocks[dev]);
            C.cached[dev] = 0;
        }
        for (auto& b : drop) (void)hipFree(b.second);
        e = hipMalloc(p, size);
    }
    if (e != hipSuccess) {
        *p = nullptr;
        return e;
    }
    std::lock_guard<std::mutex> g(C.mu);
    C.live[*p] = DevCache::Live{dev, size};
    return hipSuccess;
}

void dev_free(void* p) {
    if (!p) return;
    DevCache& C = cache();
    DevCache::Live L{-1, 0};
    {
        std::lock_guard<std::mutex> g(C.mu);
        auto it = C.live.find(p);
        if (it != C.live.end()) {
            L = it->second;
            C.live.erase(it);
        }
    }
    if (L.device < 0 || L.device >= kMaxDevices) {
        (void)hipFree(p);
        return;
    }
    int cur = 0;
    (void)hipGetDevice(&cur);
    if (cur != L.device) (void)hipSetDevice(L.device);
    const bool sync_ok = hipDeviceSynchronize() == hipSuccess;
    if (cur != L.device) (void)hipSetDevice(cur);
    bool keep = false;
    if (sync_ok) {
        std::lock_guard<std::mutex> g(C.mu);
        if (C.cached[L.device] + L.size <= kCacheCap) {
            C.free_blocks[L.device].emplace(L.size, p);
            C.cached[L.device] += L.size;
            keep = true;
        }
    }
    if (!keep) (void)hipFree(p);
}

}  // namespace saber
