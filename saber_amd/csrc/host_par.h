// Host-side parallel loop shared by the library's translation units.
#pragma once

#include <algorithm>
#include <atomic>
#include <cstdint>
#include <memory>
#include <thread>
#include <type_traits>
#include <utility>
#include <vector>

namespace saber {

// f(i) for every i in [0, n) on up to 16 host threads taking items in turn
// (on the calling thread alone below 64 items). f must be safe to run
// concurrently for different i.
template <class F>
void parallel_for(uint64_t n, F&& f) {
    const unsigned nt = std::max(1u, std::min<unsigned>(std::thread::hardware_concurrency(), 16u));
    if (n < 64 || nt == 1) {
        for (uint64_t i = 0; i < n; ++i) f(i);
        return;
    }
    std::vector<std::thread> th;
    std::atomic<uint64_t> next{0};
    for (unsigned t = 0; t < nt; ++t)
        th.emplace_back([&]() {
            for (;;) {
                const uint64_t i = next.fetch_add(1);
                if (i >= n) break;
                f(i);
            }
        });
    for (auto& x : th) x.join();
}

// An allocator whose value-less construct default-initialises: a
// host_vec<T>(n) of a trivially constructible T is not zeroed, so its pages
// are first touched by the (parallel) code that fills it.
template <class T>
struct DefaultInitAlloc : std::allocator<T> {
    template <class U>
    struct rebind {
        using other = DefaultInitAlloc<U>;
    };
    DefaultInitAlloc() = default;
    template <class U>
    DefaultInitAlloc(const DefaultInitAlloc<U>&) noexcept {}
    template <class U>
    void construct(U* p) noexcept(std::is_nothrow_default_constructible<U>::value) {
        ::new (static_cast<void*>(p)) U;
    }
    template <class U, class... A>
    void construct(U* p, A&&... a) {
        ::new (static_cast<void*>(p)) U(std::forward<A>(a)...);
    }
};
template <class T>
using host_vec = std::vector<T, DefaultInitAlloc<T>>;

// An uninitialised array of a trivially constructible T: its pages are first
// touched by the code that fills it (by several threads in a parallel fill),
// not by one thread zeroing it.
template <class T>
std::unique_ptr<T[]> raw_array(uint64_t n) {
    return std::unique_ptr<T[]>(new T[n ? n : 1]);
}

}  // namespace saber
