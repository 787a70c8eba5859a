// common.hpp -- libsqbem internals: error state, device context, RAII device buffers.
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <functional>
#include <stdexcept>
#include <string>

#include "../../include/sqbem.h"

namespace sq {

// Error carried from the internals to the C ABI boundary.
struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

void set_last_error(const std::string& msg);

#define SQ_HIP(call)                                                                        \
    do {                                                                                    \
        hipError_t e_ = (call);                                                             \
        if (e_ != hipSuccess)                                                               \
            throw ::sq::Error(SQBEM_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

#define SQ_REQUIRE(cond, code, msg)                 \
    do {                                            \
        if (!(cond)) throw ::sq::Error((code), (msg)); \
    } while (0)

// Per-process state: the device and the stream every launch goes to.
struct Context {
    int device = -1;
    hipStream_t stream = nullptr;      // library-owned stream
    hipStream_t user_stream = nullptr; // caller override (sqbem_set_stream)
    hipStream_t side = nullptr;        // second library stream (default priority)
    hipEvent_t join_ev = nullptr, fork_ev = nullptr;
    hipStream_t redirect = nullptr;    // set inside a SideScope: launches go to side
    int cu_count = 0;
    size_t hbm_bytes = 0;
    char name[256] = {0};
    hipStream_t s() const { return redirect ? redirect : (user_stream ? user_stream : stream); }
    // the library stream, also inside a SideScope
    hipStream_t lib() const { return user_stream ? user_stream : stream; }
};

Context& ctx();          // initialises device 0 on first use

// Work left on the side stream for later (tiled factor copies): side_defer starts the side
// stream after what the library stream holds and marks it pending; side_join makes the
// library stream wait for it if pending; side_drain waits on the host (before memory the
// side stream may still touch is released).
void side_defer();
void side_join();
void side_drain();

// Work issued inside this scope goes to the side stream, beside what the library stream
// holds (it must not depend on that work); when the scope ends the library stream waits
// for it. Not nested.
struct SideScope {
    SideScope();
    ~SideScope();
    SideScope(const SideScope&) = delete;
    SideScope& operator=(const SideScope&) = delete;
};
void ensure_init();      // throws on failure
void upload_tables();    // quadrature tables to constant memory (assemble.hip)
// device address of the no-allocation GEMM's scratch on this device (dense.hip), looked up
// once per device with that device current
void gemm16_scratch_setup(int device);
// Zero device bytes with a kernel on the library stream (no memset nodes, so the
// sequence stays safe inside captured graphs).
void fill_zero(void* p, size_t bytes);

// Device memory through a cache: a released block is kept (by its rounded size) and handed
// to the next request of that size instead of going back to the driver, so repeated solves
// of one size allocate nothing after the first. Library work is ordered on one stream,
// which makes a block's reuse safe without synchronisation; work issued on the side stream
// (SideScope) only receives blocks released before the library stream last drained, since
// kernels still queued there may use a block the host has released. A failed allocation
// empties the cache and retries once.
void* dev_alloc(size_t bytes);  // nullptr when the device is out of memory
void dev_free(void* p);
// the library stream has drained: every block released so far may go to side-stream work
void dev_cache_settled();
void dev_cache_release();       // return every cached block to the driver
size_t dev_cache_bytes();

// Owned device allocation (no copy, movable).
template <typename T>
struct DBuf {
    T* p = nullptr;
    size_t n = 0;
    DBuf() = default;
    explicit DBuf(size_t count) { alloc(count); }
    DBuf(const DBuf&) = delete;
    DBuf& operator=(const DBuf&) = delete;
    DBuf(DBuf&& o) noexcept : p(o.p), n(o.n) { o.p = nullptr; o.n = 0; }
    DBuf& operator=(DBuf&& o) noexcept {
        if (this != &o) { release(); p = o.p; n = o.n; o.p = nullptr; o.n = 0; }
        return *this;
    }
    ~DBuf() { release(); }
    void alloc(size_t count) {
        release();
        if (count == 0) return;
        p = (T*)dev_alloc(count * sizeof(T));
        if (!p) throw Error(SQBEM_ERR_NOMEM, "device allocation of " + std::to_string(count * sizeof(T)) + " bytes failed");
        n = count;
    }
    void release() {
        if (p) dev_free(p);
        p = nullptr;
        n = 0;
    }
    void zero() { if (p) fill_zero(p, n * sizeof(T)); }
    void upload(const T* h, size_t count) {
        SQ_HIP(hipMemcpyAsync(p, h, count * sizeof(T), hipMemcpyHostToDevice, ctx().s()));
    }
    void download(T* h, size_t count) const {
        SQ_HIP(hipMemcpyAsync(h, p, count * sizeof(T), hipMemcpyDeviceToHost, ctx().s()));
        SQ_HIP(hipStreamSynchronize(ctx().s()));
    }
    T* get() const { return p; }
};

// Sticky device error word (zero-initialised, device memory): kernels with bounded waits
// set it when a wait times out. sync() checks it whenever such a kernel was launched since
// the last check, so a timed-out wait surfaces as an error at the next synchronisation.
int* device_error_word();
void mark_device_error_check();
void check_device_error();
// zeroes the chained triangular solve's block counters (after an error word was seen)
void trsv_reset_counters();
inline void sync() {
    SQ_HIP(hipStreamSynchronize(ctx().s()));
    if (!ctx().redirect) dev_cache_settled();
    check_device_error();
}

// Graph capture of a launch sequence. While capturing() is set, work issued on the library
// stream is recorded, not run: nothing in it may synchronise, copy from host memory or read
// results back, and the 16-bit GEMM autotuner takes decided tiles without timing any.
bool& capturing();
// Runs fn (launches on the library stream) captured into a HIP graph, launched once; *exec
// receives the executable graph, which the caller destroys once the stream has passed it.
// When the capture fails or records anything but kernel launches (a memset or copy node, as
// a library call may add), the graph is discarded and fn runs directly: returns false, *exec
// null. The results are the same either way.
bool launch_as_graph(const std::function<void()>& fn, hipGraphExec_t* exec);
struct GraphExec {
    hipGraphExec_t x = nullptr;
    GraphExec() = default;
    GraphExec(const GraphExec&) = delete;
    GraphExec& operator=(const GraphExec&) = delete;
    ~GraphExec() {
        if (x) {
            (void)hipStreamSynchronize(ctx().s());
            (void)hipGraphExecDestroy(x);
        }
    }
};

// Kernel timing for roofline reports: when enabled, a ProfScope records HIP events around
// one launch on the library stream together with its algorithmic flops and bytes.
enum ProfClass { PROF_DGEMM = 0, PROF_SGEMM, PROF_GEMM16, PROF_PANEL, PROF_TRSM, PROF_ASSEMBLE, PROF_GEMV, PROF_LASWP,
                 PROF_GEMM16Q, PROF_TRSV,
                 // gemm16q by autotuner candidate 0..13 (each launch also counts in PROF_GEMM16Q)
                 PROF_G16Q_C0, PROF_G16Q_C1, PROF_G16Q_C2, PROF_G16Q_C3, PROF_G16Q_C4, PROF_G16Q_C5, PROF_G16Q_C6,
                 PROF_G16Q_C7, PROF_G16Q_C8, PROF_G16Q_C9, PROF_G16Q_C10, PROF_G16Q_C11, PROF_G16Q_C12,
                 PROF_G16Q_C13, PROF_NCLASS };
bool prof_enabled();
struct ProfScope {
    int cls;
    double flops, bytes;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    const char* name;  // kernel name (a string that outlives the process's profiling), or null
    ProfScope(int c, double f, double b, const char* kernel = nullptr);
    ~ProfScope();
};

inline void check_launch(const char* what) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) throw Error(SQBEM_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

inline unsigned ceil_div(size_t a, size_t b) { return (unsigned)((a + b - 1) / b); }

// Wrap a C ABI body: exceptions become status codes plus sqbem_last_error().
template <typename F>
int guarded(F&& f) {
    try {
        ensure_init();
        f();
        return SQBEM_OK;
    } catch (const Error& e) {
        set_last_error(e.what());
        return e.code;
    } catch (const std::bad_alloc&) {
        set_last_error("host allocation failed");
        return SQBEM_ERR_NOMEM;
    } catch (const std::exception& e) {
        set_last_error(e.what());
        return SQBEM_ERR_ARG;
    }
}

template <typename T, typename F>
T* guarded_ptr(F&& f) {
    try {
        ensure_init();
        return f();
    } catch (const std::exception& e) {
        set_last_error(e.what());
        return nullptr;
    }
}

}  // namespace sq
