// fft_any.hip -- batched 1D passes for transform lengths without a compiled
// plan.  The reference plans FFTW for any N (src/FFT1D.cpp:14-54) and its
// drivers take N from the command line (exec/ns.cpp:38-40, exec/fft.cpp:19-22);
// this file gives the passes of fft3d.hip the same reach:
//
//  * mixed radix (k_fft_any): N = product of 8, 4, 2, 3, 5, 7, 11, 13, run as
//    a Stockham autosort whose radix list is a kernel argument.  A block holds
//    LPB whole lines in LDS (2 x LPB x N complex values, ping-pong between
//    stages), every stage's butterflies spread over the block's lanes, one
//    barrier per stage.  Loads and stores use the compiled passes' lane
//    orders: element-fastest for contiguous lines, line(column)-fastest for
//    strided ones, so the global accesses stay coalesced.  N <= 2048.
//  * Bluestein (chirp-z) for the other lengths N <= 1024 (a prime factor
//    above 13): with w_k = exp(s i pi k^2 / N),
//      X_q = w_q sum_n (x_n w_n) conj(w_{q-n}),
//    a cyclic convolution of length M >= 2N - 1 (M a product of 2, 3, 5, 7)
//    done as forward FFT_M, a product with the precomputed FFT_M of the
//    chirp filter (scaled by 1/M), inverse FFT_M.  The chirp angles reduce
//    k^2 mod 2N exactly and the filter spectrum is a direct DFT summed in
//    long double on the host, once per (N, sign, precision, device).
//
// Both read and write each line once in place (per chunk, through a scratch
// buffer of M values per line for Bluestein).
#include <cmath>
#include <map>
#include <mutex>
#include <tuple>
#include <vector>

#include "fft_core.hpp"

namespace sh {

static constexpr int kAnyBlock = 256;
static constexpr int kAnyMaxN = 2048;   // 2 x N complex128 values fit 64 KiB of LDS
static constexpr int kBluMaxN = 1024;   // Bluestein's M <= 2048
static constexpr int kAnyMaxStages = 24;

struct AnyPlan {
    int n, nst;
    int radix[kAnyMaxStages];
};

// radices 8, 4, 2, 3, 5, 7, 11, 13; false if another prime factor is left
static bool any_factor(int n, AnyPlan& p) {
    p.n = n;
    p.nst = 0;
    if (n < 1) return false;
    int m = n;
    const int rs[] = {8, 4, 2, 3, 5, 7, 11, 13};
    for (int r : rs)
        while (m % r == 0 && p.nst < kAnyMaxStages) {
            p.radix[p.nst++] = r;
            m /= r;
        }
    return m == 1;
}

static bool bluestein_len(int n, int* M) {
    for (int m = 2 * n - 1; m <= kAnyMaxN; m++) {
        AnyPlan p;
        int q = m;
        for (int r : {2, 3, 5, 7})
            while (q % r == 0) q /= r;
        if (q == 1 && any_factor(m, p)) {
            *M = m;
            return true;
        }
    }
    return false;
}

bool fft_length_supported(int n) {
    if (fft_size_supported(n)) return true;
    AnyPlan p;
    if (n <= kAnyMaxN && any_factor(n, p)) return true;
    int M;
    return n >= 1 && n <= kBluMaxN && bluestein_len(n, &M);
}

// ---------------------------------------------------------------- mixed-radix kernel
// one butterfly of a compiled radix R: inputs j + r m, twiddled by
// tw[k r step] (conjugated for SIGN > 0), outputs at out[r Ns]
template <typename T, int R, int SIGN>
__device__ __forceinline__ void any_bfly(const C2<T>* in, C2<T>* out, int j, int m, int k, int step, int Ns,
                                         const C2<T>* __restrict__ tw) {
    C2<T> v[R];
#pragma unroll
    for (int r = 0; r < R; r++) v[r] = in[j + r * m];
    if (Ns > 1) {
#pragma unroll
        for (int r = 1; r < R; r++) {
            C2<T> w = tw[k * r * step];
            if (SIGN > 0) w.y = -w.y;
            v[r] = cmul<T>(v[r], w);
        }
    }
    Dft<T, R, SIGN>::run(v);
#pragma unroll
    for (int r = 0; r < R; r++) out[r * Ns] = v[r];
}

// a prime radix R (11, 13) as a direct DFT with the constants from the
// length-n table: exp(SIGN 2 pi i t / R) = tw[t n / R] (conjugated for SIGN > 0)
template <typename T, int SIGN>
__device__ __forceinline__ void any_bfly_direct(const C2<T>* in, C2<T>* out, int j, int m, int k, int step,
                                                int Ns, int R, int n, const C2<T>* __restrict__ tw) {
    const int cstep = n / R;
    for (int q = 0; q < R; q++) {
        T ar = T(0), ai = T(0);
        for (int r = 0; r < R; r++) {
            C2<T> x = in[j + r * m];
            if (Ns > 1 && r > 0) {
                C2<T> w = tw[k * r * step];
                if (SIGN > 0) w.y = -w.y;
                x = cmul<T>(x, w);
            }
            C2<T> c = tw[((r * q) % R) * cstep];
            if (SIGN > 0) c.y = -c.y;
            ar += x.x * c.x - x.y * c.y;
            ai += x.x * c.y + x.y * c.x;
        }
        out[q * Ns] = mk<T>(ar, ai);
    }
}

// line L of the batch starts at (L % a_n) a_d + (L / a_n) b_d and holds n
// values at stride es; field blockIdx.y is fstride further
struct AnyAddr {
    long nlines, a_n, a_d, b_d, es, fstride;
    int strided;
    __device__ __forceinline__ long base(long L) const { return (L % a_n) * a_d + (L / a_n) * b_d; }
};

template <typename T>
__device__ __forceinline__ C2<T> any_epilogue(C2<T> v, long L, int pos, int n, const AnyAddr& ad,
                                              const Epi<T>& ep) {
    if (ep.flags & 1) v = rmul<T>(ep.scale, v);
    if (ep.flags & 2) {
        long a, b;
        if (ad.strided) {
            a = L % ad.a_n;
            b = L / ad.a_n;
        } else {
            a = L % ep.a_n;
            b = L / ep.a_n;
        }
        if (a == ep.za || b == ep.zb || pos == n / 2) v = mk<T>(T(0), T(0));
    }
    return v;
}

template <typename T, int SIGN>
__global__ __launch_bounds__(kAnyBlock) void k_fft_any(C2<T>* __restrict__ data, const C2<T>* __restrict__ tw,
                                                       AnyPlan pl, int lpb, AnyAddr ad, Epi<T> ep) {
    extern __shared__ __align__(16) unsigned char any_smem[];
    const int n = pl.n;
    C2<T>* A = reinterpret_cast<C2<T>*>(any_smem);
    C2<T>* B = A + lpb * n;
    data += (long)blockIdx.y * ad.fstride;
    const long l0 = (long)blockIdx.x * lpb;
    const int E = lpb * n;
    for (int e = threadIdx.x; e < E; e += blockDim.x) {
        const int l = ad.strided ? e % lpb : e / n;
        const int pos = ad.strided ? e / lpb : e % n;
        const long L = l0 + l;
        A[l * n + pos] = L < ad.nlines ? data[ad.base(L) + pos * ad.es] : mk<T>(T(0), T(0));
    }
    __syncthreads();
    int Ns = 1;
    for (int s = 0; s < pl.nst; s++) {
        const int R = pl.radix[s];
        const int m = n / R;
        const int step = n / (Ns * R);
        for (int b = threadIdx.x; b < lpb * m; b += blockDim.x) {
            const int l = b / m, j = b % m, k = j % Ns;
            const C2<T>* in = A + l * n;
            C2<T>* out = B + l * n + (j / Ns) * Ns * R + k;
            switch (R) {
                case 2: any_bfly<T, 2, SIGN>(in, out, j, m, k, step, Ns, tw); break;
                case 3: any_bfly<T, 3, SIGN>(in, out, j, m, k, step, Ns, tw); break;
                case 4: any_bfly<T, 4, SIGN>(in, out, j, m, k, step, Ns, tw); break;
                case 5: any_bfly<T, 5, SIGN>(in, out, j, m, k, step, Ns, tw); break;
                case 7: any_bfly<T, 7, SIGN>(in, out, j, m, k, step, Ns, tw); break;
                case 8: any_bfly<T, 8, SIGN>(in, out, j, m, k, step, Ns, tw); break;
                default: any_bfly_direct<T, SIGN>(in, out, j, m, k, step, Ns, R, n, tw); break;
            }
        }
        __syncthreads();
        C2<T>* t = A;
        A = B;
        B = t;
        Ns *= R;
    }
    for (int e = threadIdx.x; e < E; e += blockDim.x) {
        const int l = ad.strided ? e % lpb : e / n;
        const int pos = ad.strided ? e / lpb : e % n;
        const long L = l0 + l;
        if (L < ad.nlines) data[ad.base(L) + pos * ad.es] = any_epilogue<T>(A[l * n + pos], L, pos, n, ad, ep);
    }
}

template <typename T>
static int launch_any(C2<T>* data, const AnyPlan& pl, const AnyAddr& ad, int count, int dir, const Epi<T>& ep,
                      hipStream_t st) {
    const C2<T>* tw = (const C2<T>*)twiddles(pl.n, sizeof(T) == 8 ? SH_F64 : SH_F32);
    if (!tw) {
        set_error("twiddle table allocation failed");
        return SH_ENOMEM;
    }
    int lpb = kAnyMaxN / pl.n;
    if (lpb > 64) lpb = 64;
    if (lpb < 1) lpb = 1;
    const long blocks = (ad.nlines + lpb - 1) / lpb;
    if (blocks > 0x7fffffffL || count > 65535) {
        set_error("fft: batch too large for the mixed-radix pass");
        return SH_EINVAL;
    }
    const size_t lds = 2 * (size_t)lpb * pl.n * sizeof(C2<T>);
    dim3 grid((unsigned)blocks, (unsigned)count);
    if (dir < 0)
        hipLaunchKernelGGL((k_fft_any<T, -1>), grid, dim3(kAnyBlock), lds, st, data, tw, pl, lpb, ad, ep);
    else
        hipLaunchKernelGGL((k_fft_any<T, 1>), grid, dim3(kAnyBlock), lds, st, data, tw, pl, lpb, ad, ep);
    SH_LAUNCH_CHECK("k_fft_any");
    return SH_OK;
}

// ---------------------------------------------------------------- Bluestein
// tmp line t (flattened batch line g = g0 + t: field g / nlines, line g % nlines):
// x_k w_k for k < n, zero up to M
template <typename T>
__global__ void k_blu_pre(const C2<T>* __restrict__ data, C2<T>* __restrict__ tmp, const C2<T>* __restrict__ wch,
                          int n, int M, long g0, long lc, AnyAddr ad) {
    const long total = lc * M;
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < total; i += (long)gridDim.x * blockDim.x) {
        const long t = i / M;
        const int k = (int)(i % M);
        C2<T> v = mk<T>(T(0), T(0));
        if (k < n) {
            const long g = g0 + t;
            const long f = g / ad.nlines, L = g % ad.nlines;
            v = cmul<T>(data[f * ad.fstride + ad.base(L) + k * ad.es], wch[k]);
        }
        tmp[i] = v;
    }
}

template <typename T>
__global__ void k_blu_mul(C2<T>* __restrict__ tmp, const C2<T>* __restrict__ bf, int M, long total) {
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < total; i += (long)gridDim.x * blockDim.x)
        tmp[i] = cmul<T>(tmp[i], bf[i % M]);
}

template <typename T>
__global__ void k_blu_post(C2<T>* __restrict__ data, const C2<T>* __restrict__ tmp, const C2<T>* __restrict__ wch,
                           int n, int M, long g0, long lc, AnyAddr ad, Epi<T> ep) {
    const long total = lc * n;
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < total; i += (long)gridDim.x * blockDim.x) {
        const long t = i / n;
        const int q = (int)(i % n);
        const long g = g0 + t;
        const long f = g / ad.nlines, L = g % ad.nlines;
        data[f * ad.fstride + ad.base(L) + q * ad.es] =
            any_epilogue<T>(cmul<T>(tmp[t * M + q], wch[q]), L, q, n, ad, ep);
    }
}

struct BluKey {
    int dev, n, sign, prec;
    bool operator<(const BluKey& o) const {
        return std::tie(dev, n, sign, prec) < std::tie(o.dev, o.n, o.sign, o.prec);
    }
};
struct BluTables {
    int M;
    void* wch;  // n chirp values w_k
    void* bf;   // M values: FFT_M of the chirp filter, divided by M
};
static std::mutex g_blu_mu;
static std::map<BluKey, BluTables> g_blu;

template <typename T>
static int blu_tables(int n, int sign, BluTables* out) {
    const int prec = sizeof(T) == 8 ? SH_F64 : SH_F32;
    const BluKey key{current_device(), n, sign, prec};
    {
        std::lock_guard<std::mutex> lk(g_blu_mu);
        auto it = g_blu.find(key);
        if (it != g_blu.end()) {
            *out = it->second;
            return SH_OK;
        }
    }
    int M = 0;
    if (!bluestein_len(n, &M)) {
        set_error("fft: no Bluestein length for " + std::to_string(n));
        return SH_EUNSUPPORTED;
    }
    const long double pi = 3.141592653589793238462643383279502884L;
    std::vector<long double> wr(n), wi(n);
    for (int k = 0; k < n; k++) {
        const long long q = ((long long)k * k) % (2LL * n);
        const long double a = sign * pi * (long double)q / (long double)n;
        wr[k] = cosl(a);
        wi[k] = sinl(a);
    }
    // filter b_m = conj(w_|m|) on the cyclic indices m and M - m, m < n
    std::vector<long double> br(M, 0.0L), bi(M, 0.0L);
    for (int m = 0; m < n; m++) {
        br[m] = wr[m];
        bi[m] = -wi[m];
        if (m > 0) {
            br[M - m] = wr[m];
            bi[M - m] = -wi[m];
        }
    }
    std::vector<long double> cr(M), ci(M);
    for (int t = 0; t < M; t++) {
        const long double a = 2.0L * pi * (long double)t / (long double)M;
        cr[t] = cosl(a);
        ci[t] = -sinl(a);
    }
    std::vector<T> hw(2 * (size_t)n), hb(2 * (size_t)M);
    for (int k = 0; k < n; k++) {
        hw[2 * k] = (T)wr[k];
        hw[2 * k + 1] = (T)wi[k];
    }
    for (int q = 0; q < M; q++) {
        long double sr = 0, si = 0;
        for (int m = 0; m < M; m++) {
            if (br[m] == 0 && bi[m] == 0) continue;
            const int t = (int)(((long long)m * q) % M);
            sr += br[m] * cr[t] - bi[m] * ci[t];
            si += br[m] * ci[t] + bi[m] * cr[t];
        }
        hb[2 * q] = (T)(sr / M);
        hb[2 * q + 1] = (T)(si / M);
    }
    BluTables tb{M, nullptr, nullptr};
    if (hipMalloc(&tb.wch, hw.size() * sizeof(T)) != hipSuccess ||
        hipMalloc(&tb.bf, hb.size() * sizeof(T)) != hipSuccess) {
        set_error("fft: Bluestein table allocation failed");
        return SH_ENOMEM;
    }
    SH_HIP(hipMemcpy(tb.wch, hw.data(), hw.size() * sizeof(T), hipMemcpyHostToDevice));
    SH_HIP(hipMemcpy(tb.bf, hb.data(), hb.size() * sizeof(T), hipMemcpyHostToDevice));
    std::lock_guard<std::mutex> lk(g_blu_mu);
    g_blu[key] = tb;
    *out = tb;
    return SH_OK;
}

static unsigned any_grid(long n) {
    long g = (n + 255) / 256;
    if (g > 65536L * 4) g = 65536L * 4;
    return (unsigned)(g < 1 ? 1 : g);
}

template <typename T>
static int bluestein(C2<T>* data, int n, const AnyAddr& ad, int count, int dir, const Epi<T>& ep,
                     hipStream_t st) {
    BluTables tb;
    int rc = blu_tables<T>(n, dir, &tb);
    if (rc) return rc;
    const int M = tb.M;
    const long lines = ad.nlines * count;
    long lc = (64L << 20) / ((long)M * (long)sizeof(C2<T>));  // 64 MiB of scratch per chunk
    if (lc < 1) lc = 1;
    if (lc > lines) lc = lines;
    C2<T>* tmp = nullptr;
    rc = pool_alloc((void**)&tmp, (size_t)lc * M * sizeof(C2<T>));
    if (rc) return rc;
    const Epi<T> plain{T(1), 0, 1, -1, -1};
    for (long g0 = 0; g0 < lines && !rc; g0 += lc) {
        const long c = lines - g0 < lc ? lines - g0 : lc;
        hipLaunchKernelGGL(k_blu_pre<T>, dim3(any_grid(c * M)), dim3(256), 0, st, (const C2<T>*)data, tmp,
                           (const C2<T>*)tb.wch, n, M, g0, c, ad);
        if (hipGetLastError() != hipSuccess) rc = SH_EHIP;
        if (!rc) rc = pass_contig<T>(tmp, M, c, M, 1, 0, -1, plain, st, -1);
        if (!rc) {
            hipLaunchKernelGGL(k_blu_mul<T>, dim3(any_grid(c * M)), dim3(256), 0, st, tmp, (const C2<T>*)tb.bf, M,
                               c * M);
            if (hipGetLastError() != hipSuccess) rc = SH_EHIP;
        }
        if (!rc) rc = pass_contig<T>(tmp, M, c, M, 1, 0, 1, plain, st, -1);
        if (!rc) {
            hipLaunchKernelGGL(k_blu_post<T>, dim3(any_grid(c * n)), dim3(256), 0, st, data, (const C2<T>*)tmp,
                               (const C2<T>*)tb.wch, n, M, g0, c, ad, ep);
            if (hipGetLastError() != hipSuccess) rc = SH_EHIP;
        }
    }
    if (rc == SH_EHIP) set_error("fft: Bluestein launch failed");
    pool_free(tmp);
    return rc;
}

// ---------------------------------------------------------------- entry from the passes
template <typename T>
int pass_any(C2<T>* data, int n, long nlines, long a_n, long a_d, long b_d, long es, int count, long fstride,
             int dir, const Epi<T>& ep, bool strided, hipStream_t st) {
    const AnyAddr ad{nlines, a_n, a_d, b_d, es, fstride, strided ? 1 : 0};
    if (nlines == 0 || count == 0) return SH_OK;
    AnyPlan pl;
    if (n <= kAnyMaxN && any_factor(n, pl)) return launch_any<T>(data, pl, ad, count, dir, ep, st);
    int M;
    if (n <= kBluMaxN && bluestein_len(n, &M)) return bluestein<T>(data, n, ad, count, dir, ep, st);
    set_error("unsupported FFT length " + std::to_string(n) + " (mixed radix up to " + std::to_string(kAnyMaxN) +
              ", Bluestein up to " + std::to_string(kBluMaxN) + ")");
    return SH_EUNSUPPORTED;
}

template int pass_any<double>(double2*, int, long, long, long, long, long, int, long, int, const Epi<double>&, bool,
                              hipStream_t);
template int pass_any<float>(float2*, int, long, long, long, long, long, int, long, int, const Epi<float>&, bool,
                             hipStream_t);

}  // namespace sh
