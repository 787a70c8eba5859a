// runtime.hip -- device selection, library stream, error reporting, small utility kernels.
#include <cstring>
#include <map>
#include <string>
#include <mutex>
#include <unordered_map>
#include <vector>

#include "common.hpp"

namespace sq {

namespace {
thread_local std::string g_last_error;
Context g_ctx;
std::mutex g_init_mutex;

__global__ void fill_zero_kernel(uint4* p16, size_t n16, unsigned char* tail, size_t ntail) {
    size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t k = i; k < n16; k += stride) p16[k] = make_uint4(0, 0, 0, 0);
    if (i < ntail) tail[i] = 0;
}
}  // namespace

void set_last_error(const std::string& msg) { g_last_error = msg; }

static void init_device(int device) {
    int n = 0;
    SQ_HIP(hipGetDeviceCount(&n));
    SQ_REQUIRE(n > 0, SQBEM_ERR_HIP, "no HIP device visible");
    SQ_REQUIRE(device >= 0 && device < n, SQBEM_ERR_ARG, "device index out of range");
    SQ_HIP(hipSetDevice(device));
    hipDeviceProp_t prop;
    SQ_HIP(hipGetDeviceProperties(&prop, device));
    g_ctx.device = device;
    g_ctx.cu_count = prop.multiProcessorCount;
    g_ctx.hbm_bytes = prop.totalGlobalMem;
    std::snprintf(g_ctx.name, sizeof g_ctx.name, "%s (%s)", prop.name, prop.gcnArchName);
    if (!g_ctx.stream) SQ_HIP(hipStreamCreateWithFlags(&g_ctx.stream, hipStreamNonBlocking));
    if (!g_ctx.side) {
        // The side stream (a swimmer's coupling assembly beside the tube's factorisation,
        // tiled factor copies) runs on three quarters of the CUs. The factorisation's
        // diagonal-block kernel takes a CU's whole register file, so it waits for a CU to
        // empty; with the side stream's workgroups on every CU it would wait at each of its
        // hundreds of launches. Without a CU mask the side stream is an ordinary one.
        const int cus = g_ctx.cu_count;
        std::vector<uint32_t> mask((size_t)(cus + 31) / 32, 0u);
        for (int i = 0; i < cus - cus / 4; ++i) mask[(size_t)i / 32] |= 1u << (i % 32);
        if (cus < 16 || hipExtStreamCreateWithCUMask(&g_ctx.side, (uint32_t)mask.size(), mask.data()) != hipSuccess) {
            (void)hipGetLastError();
            g_ctx.side = nullptr;
            SQ_HIP(hipStreamCreateWithFlags(&g_ctx.side, hipStreamNonBlocking));
        }
    }
    if (!g_ctx.join_ev) SQ_HIP(hipEventCreateWithFlags(&g_ctx.join_ev, hipEventDisableTiming));
    if (!g_ctx.fork_ev) SQ_HIP(hipEventCreateWithFlags(&g_ctx.fork_ev, hipEventDisableTiming));
    upload_tables();
    gemm16_scratch_setup(device);
}

void ensure_init() {
    if (g_ctx.device >= 0) {
        int cur = -1;
        SQ_HIP(hipGetDevice(&cur));
        if (cur != g_ctx.device) SQ_HIP(hipSetDevice(g_ctx.device));
        return;
    }
    std::lock_guard<std::mutex> lock(g_init_mutex);
    if (g_ctx.device < 0) init_device(0);
}

Context& ctx() { return g_ctx; }

static bool g_side_pending = false;

void side_defer() {
    SQ_HIP(hipEventRecord(g_ctx.fork_ev, g_ctx.s()));
    SQ_HIP(hipStreamWaitEvent(g_ctx.side, g_ctx.fork_ev, 0));
    g_side_pending = true;
}
void side_join() {
    if (!g_side_pending) return;
    g_side_pending = false;
    SQ_HIP(hipEventRecord(g_ctx.join_ev, g_ctx.side));
    SQ_HIP(hipStreamWaitEvent(g_ctx.s(), g_ctx.join_ev, 0));
}
void side_drain() {
    if (g_ctx.side) (void)hipStreamSynchronize(g_ctx.side);
}

SideScope::SideScope() {
    SQ_REQUIRE(!g_ctx.redirect, SQBEM_ERR_STATE, "side-stream scopes do not nest");
    g_ctx.redirect = g_ctx.side;
}
SideScope::~SideScope() {
    (void)hipEventRecord(g_ctx.join_ev, g_ctx.side);
    g_ctx.redirect = nullptr;
    (void)hipStreamWaitEvent(g_ctx.s(), g_ctx.join_ev, 0);
}

static int* g_dev_err = nullptr;
static bool g_dev_err_check = false;

int* device_error_word() {
    if (!g_dev_err) {
        SQ_HIP(hipMalloc((void**)&g_dev_err, 256));
        fill_zero(g_dev_err, 256);
    }
    return g_dev_err;
}

void mark_device_error_check() { g_dev_err_check = true; }

void check_device_error() {
    if (!g_dev_err_check || !g_dev_err) return;
    g_dev_err_check = false;
    // the chained solves that write the word run on the library stream: read it there, also
    // when called inside a side-stream scope
    int v = 0;
    SQ_HIP(hipMemcpyAsync(&v, g_dev_err, sizeof(int), hipMemcpyDeviceToHost, ctx().lib()));
    SQ_HIP(hipStreamSynchronize(ctx().lib()));
    if (v) {
        const hipStream_t saved = ctx().redirect;
        ctx().redirect = nullptr;  // the clearing kernels go to the library stream as well
        fill_zero(g_dev_err, 256);
        trsv_reset_counters();
        ctx().redirect = saved;
        SQ_HIP(hipStreamSynchronize(ctx().lib()));
        throw Error(SQBEM_ERR_HIP, v == 2 ? "triangular solve: block counter out of step with the host"
                                          : "triangular solve: a dependency wait timed out");
    }
}

// ------------------------------------------------------------------ device memory cache
namespace {
struct FreeBlock {
    void* p;
    unsigned long long seq;  // when it was released (count of releases so far)
};
struct MemCache {
    std::multimap<size_t, FreeBlock> free_blocks;  // rounded size -> block
    std::unordered_map<void*, size_t> live;        // block -> rounded size
    size_t cached = 0;
    unsigned long long released = 0;  // releases so far
    unsigned long long settled = 0;   // releases made before the library stream last drained
};
// never destroyed: buffers with static storage release into it during process exit
MemCache& mem() {
    static MemCache* m = new MemCache;
    return *m;
}
size_t round_bytes(size_t b) {
    constexpr size_t big = size_t(2) << 20;
    if (b >= big) return (b + big - 1) / big * big;
    size_t r = 256;
    while (r < b) r <<= 1;
    return r;
}
}  // namespace

void* dev_alloc(size_t bytes) {
    MemCache& M = mem();
    const size_t r = round_bytes(bytes);
    // Work on the side stream runs beside the library stream, whose queued kernels may still
    // use a block the host has already released: such a block is handed to side work only
    // once the library stream has drained after its release.
    const bool side = ctx().redirect != nullptr;
    auto range = M.free_blocks.equal_range(r);
    for (auto it = range.first; it != range.second; ++it) {
        if (side && it->second.seq > M.settled) continue;
        void* p = it->second.p;
        M.free_blocks.erase(it);
        M.cached -= r;
        M.live[p] = r;
        return p;
    }
    void* p = nullptr;
    if (hipMalloc(&p, r) != hipSuccess) {
        (void)hipGetLastError();
        dev_cache_release();
        p = nullptr;
        if (hipMalloc(&p, r) != hipSuccess) {
            (void)hipGetLastError();
            return nullptr;
        }
    }
    M.live[p] = r;
    return p;
}

void dev_free(void* p) {
    if (!p) return;
    MemCache& M = mem();
    auto it = M.live.find(p);
    if (it == M.live.end()) {
        (void)hipFree(p);
        return;
    }
    M.free_blocks.emplace(it->second, FreeBlock{p, ++M.released});
    M.cached += it->second;
    M.live.erase(it);
}

void dev_cache_settled() {
    MemCache& M = mem();
    M.settled = M.released;
}

void dev_cache_release() {
    MemCache& M = mem();
    if (M.free_blocks.empty()) return;
    (void)hipDeviceSynchronize();
    for (auto& kv : M.free_blocks) (void)hipFree(kv.second.p);
    M.free_blocks.clear();
    M.cached = 0;
}

size_t dev_cache_bytes() { return mem().cached; }

void fill_zero(void* p, size_t bytes) {
    if (!p || bytes == 0) return;
    // device allocations are 256-byte aligned; callers pass whole buffers or aligned views
    size_t n16 = bytes / 16, ntail = bytes % 16;
    unsigned blocks = (unsigned)std::min<size_t>(std::max<size_t>(ceil_div(n16, 256), 1), 8192);
    hipLaunchKernelGGL(fill_zero_kernel, dim3(blocks), dim3(256), 0, ctx().s(), (uint4*)p, n16,
                       (unsigned char*)p + 16 * n16, ntail);
    check_launch("fill_zero");
}

// ---- kernel timing
namespace {
struct ProfRec {
    int cls;
    double flops, bytes;
    hipEvent_t e0, e1;
    const char* name;
};
struct ProfTotals {
    long launches = 0;
    double flops = 0, bytes = 0, ms = 0;
};
bool g_prof = false;
unsigned g_prof_mask = 0;
std::vector<ProfRec> g_pending;
std::vector<hipEvent_t> g_event_pool;
ProfTotals g_tot[PROF_NCLASS];
std::map<std::string, ProfTotals> g_named;  // per kernel name, for launches that give one

hipEvent_t take_event() {
    if (!g_event_pool.empty()) {
        hipEvent_t e = g_event_pool.back();
        g_event_pool.pop_back();
        return e;
    }
    hipEvent_t e;
    SQ_HIP(hipEventCreate(&e));
    return e;
}
void resolve_pending() {
    for (auto& r : g_pending) {
        float ms = 0;
        SQ_HIP(hipEventSynchronize(r.e1));
        SQ_HIP(hipEventElapsedTime(&ms, r.e0, r.e1));
        for (ProfTotals* t : {&g_tot[r.cls], r.name ? &g_named[r.name] : nullptr}) {
            if (!t) continue;
            t->launches += 1;
            t->flops += r.flops;
            t->bytes += r.bytes;
            t->ms += ms;
        }
        g_event_pool.push_back(r.e0);
        g_event_pool.push_back(r.e1);
    }
    g_pending.clear();
}
}  // namespace

bool prof_enabled() { return g_prof; }

bool& capturing() {
    static thread_local bool c = false;
    return c;
}

bool launch_as_graph(const std::function<void()>& fn, hipGraphExec_t* exec) {
    *exec = nullptr;
    hipGraph_t graph = nullptr;
    bool ok = hipStreamBeginCapture(ctx().s(), hipStreamCaptureModeRelaxed) == hipSuccess;
    if (ok) {
        capturing() = true;
        try {
            fn();
        } catch (...) {
            ok = false;
        }
        capturing() = false;
        if (hipStreamEndCapture(ctx().s(), &graph) != hipSuccess) ok = false;
    }
    (void)hipGetLastError();
    if (ok && graph) {
        size_t n = 0;
        ok = hipGraphGetNodes(graph, nullptr, &n) == hipSuccess;
        std::vector<hipGraphNode_t> nodes(n);
        if (ok && n) ok = hipGraphGetNodes(graph, nodes.data(), &n) == hipSuccess;
        for (size_t i = 0; ok && i < n; ++i) {
            hipGraphNodeType t;
            ok = hipGraphNodeGetType(nodes[i], &t) == hipSuccess &&
                 (t == hipGraphNodeTypeKernel || t == hipGraphNodeTypeEmpty);
        }
    }
    if (ok && graph && hipGraphInstantiate(exec, graph, nullptr, nullptr, 0) != hipSuccess) {
        *exec = nullptr;
        ok = false;
    }
    if (graph) (void)hipGraphDestroy(graph);
    (void)hipGetLastError();
    if (!ok || !*exec) {
        if (*exec) (void)hipGraphExecDestroy(*exec);
        *exec = nullptr;
        fn();
        return false;
    }
    const hipError_t le = hipGraphLaunch(*exec, ctx().s());
    if (le != hipSuccess) {
        (void)hipGraphExecDestroy(*exec);
        *exec = nullptr;
        SQ_HIP(le);
    }
    return true;
}

ProfScope::ProfScope(int c, double f, double b, const char* kernel) : cls(c), flops(f), bytes(b), name(kernel) {
    if (c < 0 || !g_prof || !(g_prof_mask & (1u << c))) return;  // c < 0: never timed
    e0 = take_event();
    e1 = take_event();
    SQ_HIP(hipEventRecord(e0, ctx().s()));
}

ProfScope::~ProfScope() {
    if (!e0) return;
    if (hipEventRecord(e1, ctx().s()) == hipSuccess) g_pending.push_back({cls, flops, bytes, e0, e1, name});
    if (g_pending.size() > 4096) {
        try {
            resolve_pending();
        } catch (...) {
        }
    }
}

}  // namespace sq

using namespace sq;

extern "C" {

int sqbem_profile(int mask) {
    return guarded([&] {
        resolve_pending();
        for (auto& t : g_tot) t = ProfTotals{};
        g_named.clear();
        g_prof = mask != 0;
        g_prof_mask = mask == -1 ? 0xffffffffu : (unsigned)mask;
    });
}

int sqbem_profile_get(int cls, long* launches, double* flops, double* bytes, double* ms) {
    return guarded([&] {
        SQ_REQUIRE(cls >= 0 && cls < PROF_NCLASS, SQBEM_ERR_ARG, "unknown kernel class");
        resolve_pending();
        const auto& t = g_tot[cls];
        if (launches) *launches = t.launches;
        if (flops) *flops = t.flops;
        if (bytes) *bytes = t.bytes;
        if (ms) *ms = t.ms;
    });
}

int sqbem_profile_named(char* buf, int len) {
    return guarded([&] {
        SQ_REQUIRE(buf && len > 0, SQBEM_ERR_ARG, "sqbem_profile_named: no buffer");
        resolve_pending();
        std::string out;
        char line[96];
        for (const auto& kv : g_named) {
            std::snprintf(line, sizeof line, "\t%ld\t%.6e\t%.6e\t%.6f\n", kv.second.launches, kv.second.flops,
                          kv.second.bytes, kv.second.ms);
            out += kv.first + line;
        }
        SQ_REQUIRE((int)out.size() < len, SQBEM_ERR_ARG, "sqbem_profile_named: buffer too small");
        std::memcpy(buf, out.c_str(), out.size() + 1);
    });
}

const char* sqbem_last_error(void) { return g_last_error.c_str(); }

int sqbem_init(int device) {
    try {
        std::lock_guard<std::mutex> lock(g_init_mutex);
        if (g_ctx.device == device) return SQBEM_OK;
        SQ_REQUIRE(g_ctx.device < 0, SQBEM_ERR_STATE, "sqbem_init: already bound to another device");
        init_device(device);
        return SQBEM_OK;
    } catch (const Error& e) {
        set_last_error(e.what());
        return e.code;
    }
}

int sqbem_set_stream(void* s) {
    return guarded([&] {
        // cached device blocks are reused without synchronisation on the assumption that all
        // work runs on one stream: finish the old stream's work before switching
        SQ_HIP(hipStreamSynchronize(g_ctx.s()));
        g_ctx.user_stream = (hipStream_t)s;
    });
}

int sqbem_trim(double* released_gb) {
    return guarded([&] {
        const size_t b = dev_cache_bytes();
        dev_cache_release();
        if (released_gb) *released_gb = b / 1e9;
    });
}

int sqbem_sync(void) {
    return guarded([&] { sync(); });
}

int sqbem_device_info(char* name, int len, int* cu_count, double* hbm_gb) {
    return guarded([&] {
        if (name && len > 0) {
            std::strncpy(name, g_ctx.name, (size_t)len - 1);
            name[len - 1] = 0;
        }
        if (cu_count) *cu_count = g_ctx.cu_count;
        if (hbm_gb) *hbm_gb = g_ctx.hbm_bytes / 1e9;
    });
}

}  // extern "C"
