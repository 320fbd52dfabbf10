// Shared runtime helpers: error reporting through owl_last_error(), HIP
// status checks that turn into owl_status codes, and a device buffer.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstddef>
#include <cstdint>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "owl_amd.h"

namespace owl {

// Exception carrying an owl_status; caught at the C boundary.
struct Error : std::runtime_error {
    owl_status code;
    Error(owl_status c, const std::string& m) : std::runtime_error(m), code(c) {}
};

void set_last_error(const std::string& msg);
void clear_last_error();

[[noreturn]] inline void fail(owl_status code, const std::string& msg) { throw Error(code, msg); }

inline void hip_check(hipError_t e, const char* what) {
    if (e != hipSuccess) {
        if (e == hipErrorOutOfMemory)
            fail(OWL_ERR_NOMEM, std::string(what) + ": " + hipGetErrorString(e));
        fail(OWL_ERR_DEVICE, std::string(what) + ": " + hipGetErrorString(e));
    }
}
#define OWL_HIP(call) ::owl::hip_check((call), #call)

// Run f() and convert exceptions into a status (used by every entry point).
template <class F>
owl_status guarded(F&& f) {
    try {
        clear_last_error();
        f();
        return OWL_OK;
    } catch (const Error& e) {
        set_last_error(e.what());
        return e.code;
    } catch (const std::bad_alloc&) {
        set_last_error("host allocation failed");
        return OWL_ERR_NOMEM;
    } catch (const std::exception& e) {
        set_last_error(e.what());
        return OWL_ERR_ARG;
    }
}

// Device memory that DevBuf::alloc carves buffers from instead of calling
// hipMalloc, while a DeviceArenaScope of the calling thread is active: a
// bump allocator over memory the caller owns (256-byte aligned pieces, never
// freed one by one).  Running past its end fails with OWL_ERR_NOMEM.
struct DeviceArena {
    char* base = nullptr;
    size_t size = 0, used = 0;
};
inline thread_local DeviceArena* tl_arena = nullptr;

struct DeviceArenaScope {
    explicit DeviceArenaScope(DeviceArena* a) : prev(tl_arena) { tl_arena = a; }
    ~DeviceArenaScope() { tl_arena = prev; }
    DeviceArenaScope(const DeviceArenaScope&) = delete;
    DeviceArenaScope& operator=(const DeviceArenaScope&) = delete;
    DeviceArena* prev;
};

// Device allocation of n elements of T: owned (hipMalloc/hipFree), or a
// piece of the active DeviceArena (released with the arena).
template <class T>
class DevBuf {
public:
    DevBuf() = default;
    explicit DevBuf(size_t n) { alloc(n); }
    ~DevBuf() { release(); }
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    DevBuf(DevBuf&& o) noexcept : p_(o.p_), n_(o.n_), owned_(o.owned_) { o.p_ = nullptr; o.n_ = 0; }
    DevBuf& operator=(DevBuf&& o) noexcept {
        if (this != &o) { release(); p_ = o.p_; n_ = o.n_; owned_ = o.owned_; o.p_ = nullptr; o.n_ = 0; }
        return *this;
    }
    // (Re)allocate to n elements; keeps the buffer (and its contents) when the
    // size is unchanged.
    void alloc(size_t n) {
        if (p_ && n == n_) return;
        release();
        if (n == 0) return;
        if (DeviceArena* a = tl_arena) {
            const size_t at = (a->used + 255) / 256 * 256;
            if (n > (a->size - std::min(at, a->size)) / sizeof(T))
                fail(OWL_ERR_NOMEM, "the run needs more than the " + std::to_string(a->size >> 20) +
                                        " MiB of static device scratch");
            p_ = reinterpret_cast<T*>(a->base + at);
            a->used = at + n * sizeof(T);
            owned_ = false;
        } else {
            OWL_HIP(hipMalloc(&p_, n * sizeof(T)));
            owned_ = true;
        }
        n_ = n;
    }
    void release() {
        if (p_ && owned_) (void)hipFree(p_);
        p_ = nullptr;
        n_ = 0;
    }
    void upload(const T* h, size_t n, hipStream_t s) {
        if (n > n_) fail(OWL_ERR_ARG, "upload larger than buffer");
        if (n) OWL_HIP(hipMemcpyAsync(p_, h, n * sizeof(T), hipMemcpyHostToDevice, s));
    }
    void upload(const std::vector<T>& v, hipStream_t s) {
        if (v.size() > n_) alloc(v.size());
        upload(v.data(), v.size(), s);
    }
    void zero(hipStream_t s) {
        if (n_) OWL_HIP(hipMemsetAsync(p_, 0, n_ * sizeof(T), s));
    }
    T* get() const { return p_; }
    size_t size() const { return n_; }

private:
    T* p_ = nullptr;
    size_t n_ = 0;
    bool owned_ = true;
};

inline uint32_t float_bits(float f) {
    uint32_t u;
    std::memcpy(&u, &f, sizeof u);
    return u;
}

// A time loop captured once as a HIP graph and replayed while its
// signature (every device address and launch parameter its kernels take)
// is unchanged; a changed signature re-captures.  The loop must not
// synchronise while `capturing` is set.
class LoopGraph {
public:
    bool enabled = true;
    bool capturing = false;
    LoopGraph() = default;
    LoopGraph(const LoopGraph&) = delete;
    LoopGraph& operator=(const LoopGraph&) = delete;
    ~LoopGraph() { reset(); }
    void reset() {
        if (exec_) (void)hipGraphExecDestroy(exec_);
        exec_ = nullptr;
        sig_.clear();
    }
    template <class Fn>
    void run(hipStream_t st, const std::vector<uintptr_t>& sig, Fn&& loop) {
        if (exec_ && sig == sig_) {
            OWL_HIP(hipGraphLaunch(exec_, st));
            return;
        }
        reset();
        hipGraph_t g = nullptr;
        capturing = true;
        OWL_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        try {
            loop();
        } catch (...) {
            (void)hipStreamEndCapture(st, &g);
            if (g) (void)hipGraphDestroy(g);
            capturing = false;
            throw;
        }
        const hipError_t e = hipStreamEndCapture(st, &g);
        capturing = false;
        if (e != hipSuccess || !g) fail(OWL_ERR_DEVICE, "graph capture of the time loop failed");
        const hipError_t ei = hipGraphInstantiate(&exec_, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (ei != hipSuccess) {
            exec_ = nullptr;
            fail(OWL_ERR_DEVICE, "graph instantiation of the time loop failed");
        }
        sig_ = sig;
        OWL_HIP(hipGraphLaunch(exec_, st));
    }

private:
    hipGraphExec_t exec_ = nullptr;
    std::vector<uintptr_t> sig_;
};

}  // namespace owl
