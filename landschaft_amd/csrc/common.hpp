// Error handling and device memory helpers of the circuit library.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdlib>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

namespace lsc {

// std::invalid_argument maps to status 1 (ValueError in Python), every other
// failure to status 2 (RuntimeError), as the reference's binding raises them
// (src/cpp/circuit/python.cpp:67-70, grid.cpp:208).
struct HipError : std::runtime_error {
    using std::runtime_error::runtime_error;
};

inline void hip_check(hipError_t e, const char* what) {
    if (e != hipSuccess) throw HipError(std::string(what) + ": " + hipGetErrorString(e));
}
#define LSC_HIP(x) ::lsc::hip_check((x), #x)

// Device memory given by the caller (the allocation-free entry points'
// static scratch), handed out in order, 256-byte aligned. While one is
// active on the calling thread, every DevBuf of that thread takes its memory
// from it instead of hipMalloc, and gives nothing back one by one: an
// ArenaMark returns what was taken in its scope.
struct Arena {
    char* base = nullptr;
    std::size_t size = 0, used = 0;
};
inline Arena*& active_arena() {
    thread_local Arena* a = nullptr;
    return a;
}
struct ArenaMark {
    Arena* a = active_arena();
    std::size_t used = a ? a->used : 0;
    ArenaMark() = default;
    ArenaMark(const ArenaMark&) = delete;
    ArenaMark& operator=(const ArenaMark&) = delete;
    ~ArenaMark() {
        if (a) a->used = used;
    }
};
inline std::size_t arena_round(std::size_t bytes) { return (bytes + 255) & ~static_cast<std::size_t>(255); }

// A device allocation owned by the host object.
template <typename T>
struct DevBuf {
    T* p = nullptr;
    std::size_t n = 0;
    bool borrowed = false;  // from the active arena: not freed
    DevBuf() = default;
    explicit DevBuf(std::size_t count) { alloc(count); }
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    DevBuf(DevBuf&& o) noexcept : p(o.p), n(o.n), borrowed(o.borrowed) { o.p = nullptr, o.n = 0, o.borrowed = false; }
    DevBuf& operator=(DevBuf&& o) noexcept {
        if (this != &o) {
            release();
            p = o.p, n = o.n, borrowed = o.borrowed;
            o.p = nullptr, o.n = 0, o.borrowed = false;
        }
        return *this;
    }
    ~DevBuf() { release(); }
    void alloc(std::size_t count) {
        release();
        if (count == 0) return;
        if (Arena* a = active_arena()) {
            const std::size_t at = arena_round(a->used), bytes = count * sizeof(T);
            if (at + bytes > a->size)
                throw std::invalid_argument("the run needs more device memory than the static scratch of " +
                                            std::to_string(a->size >> 20) + " MiB holds (" +
                                            std::to_string((at + bytes) >> 20) + " MiB): use the allocating entry point");
            p = reinterpret_cast<T*>(a->base + at);
            a->used = at + bytes;
            borrowed = true;
        } else {
            LSC_HIP(hipMalloc(reinterpret_cast<void**>(&p), count * sizeof(T)));
        }
        n = count;
    }
    void release() {
        if (p && !borrowed) (void)hipFree(p);
        p = nullptr, n = 0, borrowed = false;
    }
    void upload(const T* h, std::size_t count, hipStream_t s) {
        LSC_HIP(hipMemcpyAsync(p, h, count * sizeof(T), hipMemcpyHostToDevice, s));
    }
    void upload(const std::vector<T>& v, hipStream_t s) {
        if (v.size() > n) alloc(v.size());
        if (!v.empty()) upload(v.data(), v.size(), s);
    }
};

template <typename T, class A>
DevBuf<T> to_device(const std::vector<T, A>& v, hipStream_t s) {
    DevBuf<T> d(v.size() ? v.size() : 1);
    if (!v.empty()) d.upload(v.data(), v.size(), s);
    return d;
}

// Kernel sequences are replayed as hipGraphs unless LSC_GRAPHS=0 (then
// launched one by one; the results are the same).
inline bool graphs_enabled() {
    if (active_arena()) return false;  // instantiating a graph may allocate
    const char* e = std::getenv("LSC_GRAPHS");
    return !(e && e[0] == '0');
}

// HIP rejects a wait on an event while the stream that recorded it is being
// captured (hipErrorStreamCaptureIsolation), even when the event was recorded
// before the capture began: captures and waits on other streams' events hold
// this lock.
inline std::mutex& capture_mutex() {
    static std::mutex m;
    return m;
}

// A fixed sequence of kernel launches, captured from a stream once and then
// replayed as one graph launch. `key` stands for everything the launches'
// arguments depend on (batch size, buffers): a new key captures again. The
// sequence may hold kernel launches only (no memset, copy or allocation).
// The legacy null stream cannot be captured: there the launches run eagerly.
struct LaunchGraph {
    hipGraphExec_t exec = nullptr;
    long long key = -1;
    LaunchGraph() = default;
    LaunchGraph(const LaunchGraph&) = delete;
    LaunchGraph& operator=(const LaunchGraph&) = delete;
    ~LaunchGraph() { clear(); }
    void clear() {
        if (exec) (void)hipGraphExecDestroy(exec);
        exec = nullptr, key = -1;
    }
    template <class F>
    void run(hipStream_t s, long long k, bool enabled, F&& launches) {
        if (!enabled || s == nullptr) {
            launches();
            return;
        }
        if (!exec || key != k) {
            clear();
            std::lock_guard<std::mutex> lk(capture_mutex());
            LSC_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
            hipGraph_t g = nullptr;
            try {
                launches();
            } catch (...) {
                (void)hipStreamEndCapture(s, &g);
                if (g) (void)hipGraphDestroy(g);
                throw;
            }
            LSC_HIP(hipStreamEndCapture(s, &g));
            hipGraphExec_t x = nullptr;
            const hipError_t e = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            LSC_HIP(e);
            exec = x, key = k;
        }
        LSC_HIP(hipGraphLaunch(exec, s));
    }
};

void set_error(const std::string& msg);

// One line on stderr per stage of a long graph solve when LSC_PROGRESS is set.
void progress(const char* fmt, ...);

template <class F>
int guarded(F&& f) {
    try {
        f();
        return 0;
    } catch (const std::invalid_argument& e) {
        set_error(e.what());
        return 1;
    } catch (const std::exception& e) {
        set_error(e.what());
        return 2;
    }
}

}  // namespace lsc
