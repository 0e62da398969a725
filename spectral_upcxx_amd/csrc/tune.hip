// tune.hip -- on-device autotuner for the radix/tile schedules.
//
// Every tunable pass role (x, y, z lines of the 3D transform; the fused z and
// Fourier kernels of the solver) has a list of compiled schedules that compute
// the same transform (fft_core.hpp, ns_step.hip).  The tuner times each one on
// the device for the caller's actual sizes and records the fastest; the
// dispatchers consult that record.  Choices are per device-agnostic
// (precision, length, role) key and persist for the process.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

#include "fft_core.hpp"

namespace sh {

struct TuneRec {
    int variant;
    double ms;
};
static std::mutex g_tmu;
static std::map<std::tuple<int, int, int>, TuneRec> g_tune;

int tune_choice(int precision, int n, int role) {
    std::lock_guard<std::mutex> lk(g_tmu);
    auto it = g_tune.find(std::make_tuple(precision, n, role));
    return it == g_tune.end() ? 0 : it->second.variant;
}

void tune_set(int precision, int n, int role, int variant, double ms) {
    std::lock_guard<std::mutex> lk(g_tmu);
    g_tune[std::make_tuple(precision, n, role)] = TuneRec{variant, ms};
}

static const char* role_name(int role) {
    static const char* names[NROLES] = {"x", "y", "z", "ns_cross_z", "ns_fourier_x", "ns_cross_zh", "ns_y_curl"};
    return role >= 0 && role < NROLES ? names[role] : "?";
}

int ns_variant_count(int precision, int n, int role);
const char* ns_variant_name(int precision, int n, int role, int v);
template <typename T>
int fft3d(void* data, int count, long fstride, int nx, int ny, int nz, int dir, int flags, void* stream);

// time `reps` forward + inverse 3D transforms of buf (after one warm pair):
// the passes run in sequence, as in use, so a schedule's effect on the next
// pass (what it leaves in the caches) counts
template <typename T>
static int time_fft3d(C2<T>* buf, int nx, int ny, int nz, int reps, hipStream_t st, hipEvent_t e0, hipEvent_t e1,
                      double* ms) {
    auto pair = [&]() -> int {
        int rc = fft3d<T>(buf, 1, 0, nx, ny, nz, -1, 0, st);
        if (!rc) rc = fft3d<T>(buf, 1, 0, nx, ny, nz, 1, 0, st);
        return rc;
    };
    int rc = pair();
    if (rc) return rc;
    SH_HIP(hipEventRecord(e0, st));
    for (int r = 0; r < reps && !rc; r++) rc = pair();
    if (rc) return rc;
    SH_HIP(hipEventRecord(e1, st));
    SH_HIP(hipEventSynchronize(e1));
    float t = 0;
    SH_HIP(hipEventElapsedTime(&t, e0, e1));
    *ms = t / (2.0 * reps);
    return SH_OK;
}

template <typename T>
static int autotune_fft3d(int nx, int ny, int nz, int reps, hipStream_t st) {
    const int prec = sizeof(T) == 8 ? SH_F64 : SH_F32;
    const long n = (long)nx * ny * nz;
    C2<T>* buf = nullptr;
    int rc = pool_alloc((void**)&buf, sizeof(C2<T>) * n);
    if (rc) return rc;
    SH_HIP(hipMemsetAsync(buf, 0, sizeof(C2<T>) * n, st));
    hipEvent_t e0, e1;
    SH_HIP(hipEventCreate(&e0));
    SH_HIP(hipEventCreate(&e1));
    const int lens[3] = {nx, ny, nz};
    // two rounds over the pass roles, each schedule timed in whole transforms
    // with the other roles at their current picks
    for (int k = 0; k < 6 && !rc; k++) {
        const int role = ROLE_X + k % 3;
        const int len = lens[role];
        const int nv = pass_variant_count(prec, len, role);
        if (nv <= 1) continue;
        std::vector<double> t(nv, INFINITY);
        auto time_variant = [&](int v) {
            tune_set(prec, len, role, v, 0);
            double ms = 0;
            rc = time_fft3d<T>(buf, nx, ny, nz, reps, st, e0, e1, &ms);
            if (!rc) t[v] = std::min(t[v], ms);
        };
        for (int v = 0; v < nv && !rc; v++) time_variant(v);
        // a playoff of the three fastest, timed alternately three more times
        // (the least time of each counts)
        std::vector<int> top(nv);
        for (int v = 0; v < nv; v++) top[v] = v;
        std::sort(top.begin(), top.end(), [&](int a, int b) { return t[a] < t[b]; });
        top.resize(std::min(nv, 3));
        for (int r = 0; r < 3 && !rc; r++)
            for (int v : top)
                if (!rc) time_variant(v);
        int best = top[0];
        for (int v : top)
            if (t[v] < t[best]) best = v;
        if (!rc) tune_set(prec, len, role, best, t[best]);
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    pool_free(buf);
    return rc;
}

}  // namespace sh

using namespace sh;

extern "C" {

int sh_autotune_fft3d(int nx, int ny, int nz, int precision, int reps, void* stream) {
    if (nx <= 0 || ny <= 0 || nz <= 0 || reps <= 0 || (precision != SH_F64 && precision != SH_F32)) {
        set_error("sh_autotune_fft3d: invalid argument");
        return SH_EINVAL;
    }
    if (!fft_size_supported(nx) || !fft_size_supported(ny) || !fft_size_supported(nz)) {
        set_error("sh_autotune_fft3d: unsupported dimensions");
        return SH_EUNSUPPORTED;
    }
    int rc = ensure_init();
    if (rc) return rc;
    return precision == SH_F64 ? autotune_fft3d<double>(nx, ny, nz, reps, as_stream(stream))
                               : autotune_fft3d<float>(nx, ny, nz, reps, as_stream(stream));
}

int sh_tune_variant_count(int precision, int n, int role, int* count) {
    if (!count || role < 0 || role >= NROLES) return SH_EINVAL;
    *count = role <= ROLE_Z ? pass_variant_count(precision, n, role) : ns_variant_count(precision, n, role);
    return SH_OK;
}

int sh_tune_set_variant(int precision, int n, int role, int variant) {
    int count = 0;
    int rc = sh_tune_variant_count(precision, n, role, &count);
    if (rc) return rc;
    if (variant < 0 || (count > 0 && variant >= count) || (count == 0 && variant != 0)) {
        set_error("sh_tune_set_variant: variant index out of range");
        return SH_EINVAL;
    }
    tune_set(precision, n, role, variant, 0);
    return SH_OK;
}

int sh_tune_reset(void) {
    std::lock_guard<std::mutex> lk(g_tmu);
    g_tune.clear();
    return SH_OK;
}

int sh_tune_report(char* buf, size_t len) {
    if (!buf || len == 0) return SH_EINVAL;
    std::string s;
    {
        std::lock_guard<std::mutex> lk(g_tmu);
        for (auto& kv : g_tune) {
            const int prec = std::get<0>(kv.first), n = std::get<1>(kv.first), role = std::get<2>(kv.first);
            const char* name = role <= ROLE_Z ? pass_variant_name(prec, n, role, kv.second.variant)
                                              : ns_variant_name(prec, n, role, kv.second.variant);
            char line[256];
            snprintf(line, sizeof(line), "fp%d n=%d %s: variant %d (%s) %.4f ms\n", prec, n,
                     role_name(role), kv.second.variant, name, kv.second.ms);
            s += line;
        }
    }
    std::strncpy(buf, s.c_str(), len - 1);
    buf[len - 1] = 0;
    return SH_OK;
}

}  // extern "C"
