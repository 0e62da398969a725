// sh_common.hpp -- shared types, complex helpers and error plumbing for the
// gfx950 spectral library.  Device code is written for CDNA4 directly:
// 64-wide wavefronts, LDS tiles, no portability layers.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

#include "../../include/spectral_hip.h"

namespace sh {

// ---------------------------------------------------------------- complex types
template <typename T> struct cvec;
template <> struct cvec<double> { using type = double2; };
template <> struct cvec<float> { using type = float2; };
template <typename T> using C2 = typename cvec<T>::type;

template <typename T>
__host__ __device__ __forceinline__ C2<T> mk(T x, T y) {
    C2<T> r;
    r.x = x;
    r.y = y;
    return r;
}
template <typename T>
__device__ __forceinline__ C2<T> cadd(C2<T> a, C2<T> b) { return mk<T>(a.x + b.x, a.y + b.y); }
template <typename T>
__device__ __forceinline__ C2<T> csub(C2<T> a, C2<T> b) { return mk<T>(a.x - b.x, a.y - b.y); }
// (a.x + i a.y)(b.x + i b.y), the std::complex operator* formula
template <typename T>
__device__ __forceinline__ C2<T> cmul(C2<T> a, C2<T> b) {
    return mk<T>(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// real scalar times complex
template <typename T>
__device__ __forceinline__ C2<T> rmul(T s, C2<T> a) { return mk<T>(s * a.x, s * a.y); }
// (0 + i f)(a.x + i a.y) = (-f a.y, f a.x): the reference's imag*f[d]*z, exact
template <typename T>
__device__ __forceinline__ C2<T> imul(T f, C2<T> a) { return mk<T>(-f * a.y, f * a.x); }

// deducible spellings for the two concrete element types
#define SH_CPLX_OVERLOADS(T)                                                                    \
    __device__ __forceinline__ C2<T> cadd(C2<T> a, C2<T> b) { return cadd<T>(a, b); }            \
    __device__ __forceinline__ C2<T> csub(C2<T> a, C2<T> b) { return csub<T>(a, b); }            \
    __device__ __forceinline__ C2<T> cmul(C2<T> a, C2<T> b) { return cmul<T>(a, b); }
SH_CPLX_OVERLOADS(double)
SH_CPLX_OVERLOADS(float)
#undef SH_CPLX_OVERLOADS

// Scalar::freq (src/Scalar.cpp:198-206): index <= N/2 -> index, else index - N
template <typename T>
__host__ __device__ __forceinline__ T freq(int ind, int n) {
    return (ind <= n / 2) ? T(ind) : T(ind - n);
}

// ---------------------------------------------------------------- errors
void set_error(const std::string& msg);
int hip_fail(hipError_t e, const char* what);

#define SH_HIP(call)                                            \
    do {                                                        \
        hipError_t _e = (call);                                 \
        if (_e != hipSuccess) return ::sh::hip_fail(_e, #call); \
    } while (0)

#define SH_LAUNCH_CHECK(name)                                       \
    do {                                                            \
        hipError_t _e = hipGetLastError();                          \
        if (_e != hipSuccess) return ::sh::hip_fail(_e, name);      \
    } while (0)

inline hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

// ---------------------------------------------------------------- runtime services (runtime.hip)
int ensure_init();
int current_device();
// FFT twiddle table exp(-2 pi i m / n), m = 0..n-1, on the current device
const void* twiddles(int n, int precision);
// pooled allocation for internal scratch
int pool_alloc(void** p, size_t bytes);
int pool_free(void* p);

// kernel-class timing (sh_prof_*)
struct ProfScope {
    int slot;
    hipStream_t stream;
    ProfScope(const char* name, hipStream_t s);
    ~ProfScope();
};

inline unsigned grid1d(long n, int block) { return (unsigned)((n + block - 1) / block); }

}  // namespace sh
