// runtime.hip -- device selection, error strings, the pooled allocator that
// stands in for upcxx::new_array (src/Scalar.cpp:95, 117), FFT twiddle tables
// and HIP-event kernel timing.
#include <cmath>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

#include "sh_common.hpp"

namespace sh {

int spectrum_scratch_init();  // ops.hip: device address of the spectrum's static scratch

static thread_local std::string g_err;

void set_error(const std::string& msg) { g_err = msg; }

int hip_fail(hipError_t e, const char* what) {
    g_err = std::string(what) + ": " + hipGetErrorString(e);
    return SH_EHIP;
}

static std::mutex g_mu;
static std::vector<bool> g_dev_init;

int current_device() {
    int d = 0;
    if (hipGetDevice(&d) != hipSuccess) return 0;
    return d;
}

int ensure_init() {
    int d = 0;
    SH_HIP(hipGetDevice(&d));
    std::lock_guard<std::mutex> lk(g_mu);
    if ((int)g_dev_init.size() <= d) g_dev_init.resize(d + 1, false);
    g_dev_init[d] = true;
    return SH_OK;
}

// ---------------------------------------------------------------- pool
struct Block {
    size_t bytes;
    int device;
};
static std::unordered_map<void*, Block> g_live;
static std::multimap<std::pair<int, size_t>, void*> g_free;  // (device, bytes) -> ptr

static size_t round_bytes(size_t b) {
    const size_t g = 256;
    return (b + g - 1) / g * g;
}

int pool_alloc(void** p, size_t bytes) {
    if (!p) {
        set_error("sh_malloc: null out pointer");
        return SH_EINVAL;
    }
    *p = nullptr;
    if (bytes == 0) return SH_OK;
    const int dev = current_device();
    const size_t rb = round_bytes(bytes);
    {
        std::lock_guard<std::mutex> lk(g_mu);
        auto it = g_free.lower_bound({dev, rb});
        if (it != g_free.end() && it->first.first == dev && it->first.second <= 2 * rb) {
            *p = it->second;
            g_live[*p] = Block{it->first.second, dev};
            g_free.erase(it);
            return SH_OK;
        }
    }
    void* q = nullptr;
    hipError_t e = hipMalloc(&q, rb);
    if (e != hipSuccess) {
        // release cached blocks and retry once
        (void)hipGetLastError();
        sh_pool_trim();
        e = hipMalloc(&q, rb);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            set_error("sh_malloc: hipMalloc of " + std::to_string(rb) + " bytes failed");
            return SH_ENOMEM;
        }
    }
    std::lock_guard<std::mutex> lk(g_mu);
    g_live[q] = Block{rb, dev};
    *p = q;
    return SH_OK;
}

int pool_free(void* p) {
    if (!p) return SH_OK;
    std::lock_guard<std::mutex> lk(g_mu);
    auto it = g_live.find(p);
    if (it == g_live.end()) {
        set_error("sh_free: pointer not allocated by sh_malloc");
        return SH_EINVAL;
    }
    g_free.insert({{it->second.device, it->second.bytes}, p});
    g_live.erase(it);
    return SH_OK;
}

// ---------------------------------------------------------------- twiddles
struct TwKey {
    int dev, n, prec;
    bool operator<(const TwKey& o) const {
        if (dev != o.dev) return dev < o.dev;
        if (n != o.n) return n < o.n;
        return prec < o.prec;
    }
};
static std::map<TwKey, void*> g_tw;

const void* twiddles(int n, int precision) {
    const int dev = current_device();
    {
        std::lock_guard<std::mutex> lk(g_mu);
        auto it = g_tw.find(TwKey{dev, n, precision});
        if (it != g_tw.end()) return it->second;
    }
    // exp(-2 pi i m / n) in extended precision, rounded once
    const long double pi = 3.141592653589793238462643383279502884L;
    std::vector<double> h(2 * (size_t)n);
    for (int m = 0; m < n; m++) {
        long double a = 2.0L * pi * (long double)m / (long double)n;
        h[2 * m] = (double)cosl(a);
        h[2 * m + 1] = (double)(-sinl(a));
    }
    void* d = nullptr;
    size_t bytes;
    if (precision == SH_F32) {
        std::vector<float> f(2 * (size_t)n);
        for (int m = 0; m < n; m++) {
            long double a = 2.0L * pi * (long double)m / (long double)n;
            f[2 * m] = (float)cosl(a);
            f[2 * m + 1] = (float)(-sinl(a));
        }
        bytes = f.size() * sizeof(float);
        if (hipMalloc(&d, bytes) != hipSuccess) return nullptr;
        if (hipMemcpy(d, f.data(), bytes, hipMemcpyHostToDevice) != hipSuccess) return nullptr;
    } else {
        bytes = h.size() * sizeof(double);
        if (hipMalloc(&d, bytes) != hipSuccess) return nullptr;
        if (hipMemcpy(d, h.data(), bytes, hipMemcpyHostToDevice) != hipSuccess) return nullptr;
    }
    std::lock_guard<std::mutex> lk(g_mu);
    g_tw[TwKey{dev, n, precision}] = d;
    return d;
}

// ---------------------------------------------------------------- kernel timing
struct ProfSlot {
    std::string name;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> ev;
    double total_ms = 0;
    long launches = 0;
};
static bool g_prof_on = false;
static std::vector<ProfSlot> g_prof;
static std::vector<hipEvent_t> g_event_pool;

static hipEvent_t take_event() {
    if (!g_event_pool.empty()) {
        hipEvent_t e = g_event_pool.back();
        g_event_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

static int prof_slot(const char* name) {
    for (size_t i = 0; i < g_prof.size(); i++)
        if (g_prof[i].name == name) return (int)i;
    g_prof.push_back(ProfSlot{});
    g_prof.back().name = name;
    return (int)g_prof.size() - 1;
}

ProfScope::ProfScope(const char* name, hipStream_t s) : slot(-1), stream(s) {
    if (!g_prof_on) return;
    std::lock_guard<std::mutex> lk(g_mu);
    slot = prof_slot(name);
    hipEvent_t a = take_event(), b = take_event();
    if (!a || !b) {
        slot = -1;
        return;
    }
    g_prof[slot].ev.push_back({a, b});
    (void)hipEventRecord(a, s);
}

ProfScope::~ProfScope() {
    if (slot < 0) return;
    std::lock_guard<std::mutex> lk(g_mu);
    (void)hipEventRecord(g_prof[slot].ev.back().second, stream);
}

static void prof_drain(ProfSlot& ps) {
    for (auto& pr : ps.ev) {
        (void)hipEventSynchronize(pr.second);
        float ms = 0;
        if (hipEventElapsedTime(&ms, pr.first, pr.second) == hipSuccess) {
            ps.total_ms += ms;
            ps.launches++;
        }
        g_event_pool.push_back(pr.first);
        g_event_pool.push_back(pr.second);
    }
    ps.ev.clear();
}

}  // namespace sh

using namespace sh;

extern "C" {

const char* sh_last_error(void) { return g_err.c_str(); }

int sh_init(int device) {
    int n = 0;
    SH_HIP(hipGetDeviceCount(&n));
    if (device < 0 || device >= n) {
        set_error("sh_init: device index out of range");
        return SH_EINVAL;
    }
    SH_HIP(hipSetDevice(device));
    int rc = ensure_init();
    if (rc) return rc;
    return spectrum_scratch_init();
}

int sh_device_count(int* count) {
    if (!count) return SH_EINVAL;
    SH_HIP(hipGetDeviceCount(count));
    return SH_OK;
}

int sh_malloc(void** ptr, size_t bytes) { return pool_alloc(ptr, bytes); }
int sh_free(void* ptr) { return pool_free(ptr); }

int sh_pool_trim(void) {
    std::vector<std::pair<int, void*>> rel;
    {
        std::lock_guard<std::mutex> lk(g_mu);
        for (auto& kv : g_free) rel.push_back({kv.first.first, kv.second});
        g_free.clear();
    }
    int cur = current_device();
    for (auto& r : rel) {
        (void)hipSetDevice(r.first);
        (void)hipFree(r.second);
    }
    (void)hipSetDevice(cur);
    return SH_OK;
}

int sh_memcpy_h2d(void* dst, const void* src, size_t bytes) {
    if (bytes == 0) return SH_OK;
    if (!dst || !src) return SH_EINVAL;
    SH_HIP(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
    return SH_OK;
}

int sh_memcpy_d2h(void* dst, const void* src, size_t bytes) {
    if (bytes == 0) return SH_OK;
    if (!dst || !src) return SH_EINVAL;
    SH_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
    return SH_OK;
}

int sh_memcpy_d2d(void* dst, const void* src, size_t bytes, void* stream) {
    if (bytes == 0) return SH_OK;
    if (!dst || !src) return SH_EINVAL;
    SH_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, as_stream(stream)));
    return SH_OK;
}

}  // extern "C"

namespace sh {
// Streaming copy, the bandwidth reference for the kernels' rooflines: every
// lane moves 16-byte vectors with nontemporal loads and stores (no cache
// allocation, as the passes that stream whole fields), 4 vectors in flight
// per lane per trip.
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
static constexpr int kCopyBlock = 256, kCopyUnroll = 4;

__global__ __launch_bounds__(kCopyBlock) void k_stream_copy(const u32x4* __restrict__ src,
                                                            u32x4* __restrict__ dst, long n) {
    const long stride = (long)gridDim.x * kCopyBlock * kCopyUnroll;
    for (long i = (long)blockIdx.x * kCopyBlock * kCopyUnroll + threadIdx.x; i < n; i += stride) {
        u32x4 v[kCopyUnroll];
#pragma unroll
        for (int u = 0; u < kCopyUnroll; u++)
            if (i + u * kCopyBlock < n) v[u] = __builtin_nontemporal_load(src + i + u * kCopyBlock);
#pragma unroll
        for (int u = 0; u < kCopyUnroll; u++)
            if (i + u * kCopyBlock < n) __builtin_nontemporal_store(v[u], dst + i + u * kCopyBlock);
    }
}
}  // namespace sh

extern "C" {

int sh_stream_copy(void* dst, const void* src, size_t bytes, int blocks, void* stream) {
    if (bytes == 0) return SH_OK;
    if (!dst || !src || bytes % 16 || ((uintptr_t)dst | (uintptr_t)src) % 16 || blocks < 0) {
        set_error("sh_stream_copy: 16-byte aligned pointers, a multiple of 16 bytes, blocks >= 0");
        return SH_EINVAL;
    }
    const long n = (long)(bytes / 16);
    const long per = (long)kCopyBlock * kCopyUnroll;
    long g = blocks > 0 ? blocks : (n + per - 1) / per;
    if (g > 0x7fffffffL) g = 0x7fffffffL;
    hipStream_t st = as_stream(stream);
    {
        ProfScope ps("stream_copy", st);
        hipLaunchKernelGGL(k_stream_copy, dim3((unsigned)g), dim3(kCopyBlock), 0, st, (const u32x4*)src,
                           (u32x4*)dst, n);
    }
    SH_LAUNCH_CHECK("k_stream_copy");
    return SH_OK;
}

int sh_memset(void* dst, int value, size_t bytes, void* stream) {
    if (bytes == 0) return SH_OK;
    if (!dst) return SH_EINVAL;
    SH_HIP(hipMemsetAsync(dst, value, bytes, as_stream(stream)));
    return SH_OK;
}

int sh_stream_synchronize(void* stream) {
    SH_HIP(hipStreamSynchronize(as_stream(stream)));
    return SH_OK;
}

int sh_prof_enable(int on) {
    std::lock_guard<std::mutex> lk(g_mu);
    g_prof_on = on != 0;
    return SH_OK;
}

int sh_prof_reset(void) {
    std::lock_guard<std::mutex> lk(g_mu);
    for (auto& ps : g_prof) {
        prof_drain(ps);
        ps.total_ms = 0;
        ps.launches = 0;
    }
    return SH_OK;
}

int sh_prof_query(const char* name, double* total_ms, long* launches) {
    if (!name) return SH_EINVAL;
    std::lock_guard<std::mutex> lk(g_mu);
    for (auto& ps : g_prof) {
        if (ps.name == name) {
            prof_drain(ps);
            if (total_ms) *total_ms = ps.total_ms;
            if (launches) *launches = ps.launches;
            return SH_OK;
        }
    }
    if (total_ms) *total_ms = 0;
    if (launches) *launches = 0;
    return SH_OK;
}

int sh_stream_create(void** stream) {
    if (!stream) return SH_EINVAL;
    hipStream_t s = nullptr;
    SH_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));  // default priority
    *stream = s;
    return SH_OK;
}

int sh_stream_destroy(void* stream) {
    if (!stream) return SH_OK;
    SH_HIP(hipStreamDestroy(as_stream(stream)));
    return SH_OK;
}

int sh_graph_begin(void* stream) {
    if (!stream) {
        set_error("sh_graph_begin: the default stream cannot be captured; use sh_stream_create");
        return SH_EINVAL;
    }
    SH_HIP(hipStreamBeginCapture(as_stream(stream), hipStreamCaptureModeThreadLocal));
    return SH_OK;
}

int sh_graph_end(void* stream, void** graph) {
    if (!stream || !graph) return SH_EINVAL;
    hipGraph_t g = nullptr;
    SH_HIP(hipStreamEndCapture(as_stream(stream), &g));
    hipGraphExec_t ge = nullptr;
    hipError_t e = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) return hip_fail(e, "hipGraphInstantiate");
    *graph = ge;
    return SH_OK;
}

int sh_graph_launch(void* graph, void* stream) {
    if (!graph) return SH_EINVAL;
    SH_HIP(hipGraphLaunch((hipGraphExec_t)graph, as_stream(stream)));
    return SH_OK;
}

int sh_graph_destroy(void* graph) {
    if (!graph) return SH_OK;
    SH_HIP(hipGraphExecDestroy((hipGraphExec_t)graph));
    return SH_OK;
}

int sh_prof_names(char* buf, size_t len) {
    if (!buf || len == 0) return SH_EINVAL;
    std::lock_guard<std::mutex> lk(g_mu);
    std::string s;
    for (auto& ps : g_prof) s += ps.name + "\n";
    std::strncpy(buf, s.c_str(), len - 1);
    buf[len - 1] = 0;
    return SH_OK;
}

}  // extern "C"
