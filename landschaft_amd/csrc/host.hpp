// Host threads of the symbolic work (graph building, component labels,
// nested dissection, plans).
//
// Every parallel pass here writes only slots that its index range owns, or
// combines per-thread results in a fixed order, so the outputs do not depend
// on the thread count: LSC_HOST_THREADS (1 to 16) sets it, by default the
// hardware's count capped at 16.
#pragma once

#include <algorithm>
#include <atomic>
#include <cstdint>
#include <cstdlib>
#include <exception>
#include <memory>
#include <mutex>
#include <new>
#include <thread>
#include <utility>
#include <vector>

#include <sys/mman.h>

namespace lsc {

// One line on stderr per stage of a long graph solve when LSC_PROGRESS is set.
void progress(const char* fmt, ...);

inline int host_threads() {
    if (const char* e = std::getenv("LSC_HOST_THREADS")) {
        const long v = std::strtol(e, nullptr, 10);
        if (v >= 1) return static_cast<int>(std::min<long>(v, 16));
    }
    return static_cast<int>(std::max(1u, std::min(16u, std::thread::hardware_concurrency())));
}

// The number of chunks host_chunks splits [0, n) into: one per 2^16 indices,
// at most one per host thread.
inline std::int64_t host_chunk_count(std::int64_t n) {
    return std::min<std::int64_t>(host_threads(), std::max<std::int64_t>(1, n >> 16));
}

// f(chunk, lo, hi) over [0, n) in host_chunk_count(n) contiguous chunks, one
// thread each (chunk 0 on the caller's thread).
template <class F>
void host_chunks(std::int64_t n, F&& f) {
    const std::int64_t nt = host_chunk_count(n);
    std::vector<std::thread> ts;
    for (std::int64_t t = 1; t < nt; ++t) ts.emplace_back([&, t] { f(t, n * t / nt, n * (t + 1) / nt); });
    f(std::int64_t(0), std::int64_t(0), n / nt);
    for (auto& th : ts) th.join();
}

// f(task) for every task in [0, n), taken in order by up to host_threads()
// threads as each finishes its last one; the first exception thrown is
// rethrown once all threads have stopped.
template <class F>
void host_tasks(std::int64_t n, F&& f) {
    const std::int64_t nt = std::min<std::int64_t>(host_threads(), n);
    std::atomic<std::int64_t> next{0};
    std::exception_ptr err;
    std::mutex m;
    auto run = [&] {
        for (std::int64_t i; (i = next.fetch_add(1)) < n;) {
            try {
                f(i);
            } catch (...) {
                std::lock_guard<std::mutex> g(m);
                if (!err) err = std::current_exception();
                next.store(n);
            }
        }
    };
    std::vector<std::thread> ts;
    for (std::int64_t t = 1; t < nt; ++t) ts.emplace_back(run);
    run();
    for (auto& th : ts) th.join();
    if (err) std::rethrow_exception(err);
}

// Inclusive prefix sums of counts[0..n) into out[1..n], out[0] = 0, over
// host chunks (per-chunk totals first, then each chunk's own scan). counts
// may be out + 1 (in place): each slot is read before it is written, by the
// same thread.
template <class T, class C>
void host_prefix(std::int64_t n, const C* counts, T* out) {
    const std::int64_t nt = host_chunk_count(n);
    std::vector<T> tot(static_cast<std::size_t>(nt) + 1, 0);
    host_chunks(n, [&](std::int64_t t, std::int64_t lo, std::int64_t hi) {
        T s = 0;
        for (std::int64_t i = lo; i < hi; ++i) s += static_cast<T>(counts[i]);
        tot[t + 1] = s;
    });
    for (std::int64_t t = 0; t < nt; ++t) tot[t + 1] += tot[t];
    out[0] = 0;
    host_chunks(n, [&](std::int64_t t, std::int64_t lo, std::int64_t hi) {
        T s = tot[t];
        for (std::int64_t i = lo; i < hi; ++i) {
            s += static_cast<T>(counts[i]);
            out[i + 1] = s;
        }
    });
}

// A vector whose resize leaves new elements uninitialised (trivial types),
// so that large arrays are first touched by the parallel passes that fill
// them rather than zeroed on one thread. Blocks of 32 MiB or more are
// 2 MiB-aligned and advised as huge pages (where the kernel grants them:
// 512 times fewer page faults on first touch).
template <class T>
struct NoInit : std::allocator<T> {
    template <class U>
    struct rebind {
        using other = NoInit<U>;
    };
    static constexpr std::size_t kHuge = std::size_t(2) << 20, kLarge = std::size_t(32) << 20;
    NoInit() = default;
    template <class U>
    NoInit(const NoInit<U>&) noexcept {}
    T* allocate(std::size_t n) {
        const std::size_t bytes = n * sizeof(T);
        if (bytes < kLarge) return std::allocator<T>::allocate(n);
        const std::size_t size = (bytes + kHuge - 1) & ~(kHuge - 1);
        void* p = std::aligned_alloc(kHuge, size);
        if (!p) throw std::bad_alloc();
        (void)madvise(p, size, MADV_HUGEPAGE);
        return static_cast<T*>(p);
    }
    void deallocate(T* p, std::size_t n) noexcept {
        if (n * sizeof(T) < kLarge) std::allocator<T>::deallocate(p, n);
        else std::free(p);
    }
    template <class U>
    void construct(U* p) noexcept {
        ::new (static_cast<void*>(p)) U;
    }
    template <class U, class... A>
    void construct(U* p, A&&... a) {
        ::new (static_cast<void*>(p)) U(std::forward<A>(a)...);
    }
};
template <class T>
using RawVec = std::vector<T, NoInit<T>>;

}  // namespace lsc
