// fft_core.hpp -- register/LDS line FFT for gfx950.
//
// A line of N complex values is owned by TPL threads.  The transform is a
// Stockham autosort over radices R0..R3 (N = R0*R1*R2*R3): in stage s each
// thread runs ceil(N/R_s / TPL) radix-R_s butterflies held in registers, and
// the data move between stages through LDS (re and im in separate arrays).
// Stage s, butterfly j (Ns = R0*...*R_{s-1}):
//     in : positions j + r*N/R_s,                       r = 0..R_s-1
//     twiddle: exp(sign 2 pi i (j mod Ns) r / (Ns R_s))  (from the N-entry table)
//     out: positions (j/Ns)*Ns*R_s + (j mod Ns) + r*Ns
// so the first stage reads, and the last stage writes, position
// tid + b*TPL + r*N/R: consecutive threads touch consecutive elements, which
// makes the global loads/stores of the contiguous pass coalesced with no
// extra LDS round trip.
//
// The in-register DFTs use exact-symmetric constants (cos/sin of 2 pi k/32
// evaluated in extended precision and rounded once); the inter-stage
// twiddles come from a host-built table (twiddles(), long double).
#pragma once

#include "sh_common.hpp"

namespace sh {

// cos(2 pi k / 32), sin(2 pi k / 32), k = 0..15
constexpr double kCos32[16] = {1.0, 0.9807852804032304, 0.9238795325112867, 0.8314696123025452,
                               0.7071067811865476, 0.5555702330196022, 0.3826834323650898,
                               0.19509032201612828, 0.0, -0.19509032201612828,
                               -0.3826834323650898, -0.5555702330196022, -0.7071067811865476,
                               -0.8314696123025452, -0.9238795325112867, -0.9807852804032304};
constexpr double kSin32[16] = {0.0, 0.19509032201612828, 0.3826834323650898, 0.5555702330196022,
                               0.7071067811865476, 0.8314696123025452, 0.9238795325112867,
                               0.9807852804032304, 1.0, 0.9807852804032304, 0.9238795325112867,
                               0.8314696123025452, 0.7071067811865476, 0.5555702330196022,
                               0.3826834323650898, 0.19509032201612828};
constexpr double kSqrtHalf = 0.7071067811865476;
constexpr double kSqrt3Half = 0.8660254037844386;
// cos, sin(2 pi m / R), m = 0..R/2, for the odd radices 5 and 7 (extended
// precision, rounded once)
constexpr double kCos5[3] = {1.0, 0.30901699437494745, -0.8090169943749475};
constexpr double kSin5[3] = {0.0, 0.9510565162951535, 0.5877852522924731};
constexpr double kCos7[4] = {1.0, 0.6234898018587335, -0.2225209339563144, -0.9009688679024191};
constexpr double kSin7[4] = {0.0, 0.7818314824680298, 0.9749279121818236, 0.4338837391175581};

template <int R>
__host__ __device__ constexpr double odd_cos(int m) {
    return R == 5 ? kCos5[m] : kCos7[m];
}
template <int R>
__host__ __device__ constexpr double odd_sin(int m) {
    return R == 5 ? kSin5[m] : kSin7[m];
}

// In-register DFT of an odd prime size R (5 or 7), SIGN = +-1:
// t_k = x_k + x_{R-k}, d_k = x_k - x_{R-k} (k = 1..R/2);
// X_0 = x_0 + sum t_k; for q = 1..R/2
//   a_q = x_0 + sum_k cos(2 pi kq/R) t_k,  b_q = SIGN sum_k sin(2 pi kq/R) d_k,
//   X_q = a_q + i b_q,  X_{R-q} = a_q - i b_q.
template <typename T, int R, int SIGN>
__device__ __forceinline__ void dft_odd(C2<T>* x) {
    constexpr int H = R / 2;
    C2<T> t[H + 1], d[H + 1];
#pragma unroll
    for (int k = 1; k <= H; k++) {
        t[k] = cadd<T>(x[k], x[R - k]);
        d[k] = csub<T>(x[k], x[R - k]);
    }
    const C2<T> x0 = x[0];
    C2<T> s0 = x0;
#pragma unroll
    for (int k = 1; k <= H; k++) s0 = cadd<T>(s0, t[k]);
#pragma unroll
    for (int q = 1; q <= H; q++) {
        T ar = x0.x, ai = x0.y, br = T(0), bi = T(0);
#pragma unroll
        for (int k = 1; k <= H; k++) {
            const int m = (k * q) % R;
            const T c = T(odd_cos<R>(m <= H ? m : R - m));
            const T sn = m <= H ? T(odd_sin<R>(m)) : -T(odd_sin<R>(R - m));
            ar += c * t[k].x;
            ai += c * t[k].y;
            br += sn * d[k].x;
            bi += sn * d[k].y;
        }
        if (SIGN < 0) {
            br = -br;
            bi = -bi;
        }
        x[q] = mk<T>(ar - bi, ai + br);
        x[R - q] = mk<T>(ar + bi, ai - br);
    }
    x[0] = s0;
}

// a * exp(SIGN 2 pi i K / R), R | 32, 0 <= K < R/2
template <typename T, int R, int K, int SIGN>
__device__ __forceinline__ C2<T> twc(C2<T> a) {
    static_assert(32 % R == 0, "radix must divide 32");
    if constexpr (K == 0) {
        return a;
    } else if constexpr (4 * K == R) {  // exp(SIGN i pi/2) = SIGN i
        return SIGN > 0 ? mk<T>(-a.y, a.x) : mk<T>(a.y, -a.x);
    } else if constexpr (8 * K == R || 8 * K == 3 * R) {
        // c = +-h, s = SIGN*h: (x + i y)(c + i s)
        constexpr T h = T(kSqrtHalf);
        constexpr int sc = (8 * K == R) ? 1 : -1;
        if constexpr (sc > 0) {
            return SIGN > 0 ? mk<T>(h * (a.x - a.y), h * (a.x + a.y))
                            : mk<T>(h * (a.x + a.y), h * (a.y - a.x));
        } else {
            return SIGN > 0 ? mk<T>(-h * (a.x + a.y), h * (a.x - a.y))
                            : mk<T>(h * (a.y - a.x), -h * (a.x + a.y));
        }
    } else {
        constexpr int m = K * (32 / R);
        constexpr T c = T(kCos32[m]);
        constexpr T s = T(SIGN) * T(kSin32[m]);
        return mk<T>(a.x * c - a.y * s, a.x * s + a.y * c);
    }
}

template <typename T, int R, int SIGN, int K>
__device__ __forceinline__ void dit_combine(C2<T>* x, const C2<T>* e, const C2<T>* o) {
    if constexpr (K < R / 2) {
        const C2<T> t = twc<T, R, K, SIGN>(o[K]);
        x[K] = cadd<T>(e[K], t);
        x[K + R / 2] = csub<T>(e[K], t);
        dit_combine<T, R, SIGN, K + 1>(x, e, o);
    }
}

// In-register DFT of size R: X[q] = sum_r x[r] exp(SIGN 2 pi i r q / R), natural order.
template <typename T, int R, int SIGN>
struct Dft {
    __device__ static __forceinline__ void run(C2<T>* x) {
        if constexpr (R == 1) {
        } else if constexpr (R == 2) {
            const C2<T> a = x[0], b = x[1];
            x[0] = cadd<T>(a, b);
            x[1] = csub<T>(a, b);
        } else if constexpr (R == 3) {
            constexpr T h = T(SIGN) * T(kSqrt3Half);
            const C2<T> t = cadd<T>(x[1], x[2]);
            const C2<T> m = mk<T>(x[0].x - T(0.5) * t.x, x[0].y - T(0.5) * t.y);
            const C2<T> e = csub<T>(x[1], x[2]);
            const C2<T> d = mk<T>(-h * e.y, h * e.x);
            x[0] = cadd<T>(x[0], t);
            x[1] = cadd<T>(m, d);
            x[2] = csub<T>(m, d);
        } else if constexpr (R == 5 || R == 7) {
            dft_odd<T, R, SIGN>(x);
        } else if constexpr (R == 4) {
            const C2<T> a = cadd<T>(x[0], x[2]), b = csub<T>(x[0], x[2]);
            const C2<T> c = cadd<T>(x[1], x[3]), e = csub<T>(x[1], x[3]);
            const C2<T> d = SIGN > 0 ? mk<T>(-e.y, e.x) : mk<T>(e.y, -e.x);
            x[0] = cadd<T>(a, c);
            x[2] = csub<T>(a, c);
            x[1] = cadd<T>(b, d);
            x[3] = csub<T>(b, d);
        } else {
            static_assert(R % 2 == 0 && 32 % R == 0, "unsupported radix");
            C2<T> e[R / 2], o[R / 2];
#pragma unroll
            for (int i = 0; i < R / 2; i++) {
                e[i] = x[2 * i];
                o[i] = x[2 * i + 1];
            }
            Dft<T, R / 2, SIGN>::run(e);
            Dft<T, R / 2, SIGN>::run(o);
            dit_combine<T, R, SIGN, 0>(x, e, o);
        }
    }
};

template <int N_, int TPL_, int R0, int R1 = 1, int R2 = 1, int R3 = 1, int R4 = 1>
struct FftPlan {
    static constexpr int N = N_;
    static constexpr int TPL = TPL_;
    static constexpr int S = (R1 == 1) ? 1 : (R2 == 1) ? 2 : (R3 == 1) ? 3 : (R4 == 1) ? 4 : 5;
    static constexpr int radix(int s) {
        return s == 0 ? R0 : s == 1 ? R1 : s == 2 ? R2 : s == 3 ? R3 : R4;
    }
    static constexpr int ns(int s) {
        int p = 1;
        for (int i = 0; i < s; i++) p *= radix(i);
        return p;
    }
    static constexpr int nbf(int s) { return N / radix(s); }
    static constexpr int nb(int s) { return (nbf(s) + TPL - 1) / TPL; }
    static constexpr bool full(int s) { return nbf(s) % TPL == 0; }
    static constexpr int slots(int s) { return nb(s) * radix(s); }
    static constexpr int emax() {
        int e = 0;
        for (int s = 0; s < S; s++) e = slots(s) > e ? slots(s) : e;
        return e;
    }
    static constexpr int E = emax();
    static_assert(R0 * R1 * R2 * R3 * R4 == N, "radices must multiply to N");
};

template <class P, int s, int SIGN, typename T>
__device__ __forceinline__ void stage_bfly(C2<T>* x, int tid, const C2<T>* __restrict__ tw) {
    constexpr int R = P::radix(s), Ns = P::ns(s), NB = P::nb(s);
#pragma unroll
    for (int b = 0; b < NB; b++) {
        const int j = tid + b * P::TPL;
        if (P::full(s) || j < P::nbf(s)) {
            C2<T>* v = x + b * R;
            if constexpr (Ns > 1) {
                const int k = j % Ns;
                constexpr int step = P::N / (Ns * R);
#pragma unroll
                for (int r = 1; r < R; r++) {
                    C2<T> w = tw[k * r * step];
                    if constexpr (SIGN > 0) w.y = -w.y;
                    v[r] = cmul<T>(v[r], w);
                }
            }
            Dft<T, R, SIGN>::run(v);
        }
    }
}

template <class P, int s, typename T, class Idx>
__device__ __forceinline__ void lds_put(const C2<T>* x, int tid, T* re, T* im, const Idx& idx) {
    constexpr int R = P::radix(s), Ns = P::ns(s), NB = P::nb(s);
#pragma unroll
    for (int b = 0; b < NB; b++) {
        const int j = tid + b * P::TPL;
        if (P::full(s) || j < P::nbf(s)) {
            const int base = (j / Ns) * Ns * R + (j % Ns);
#pragma unroll
            for (int r = 0; r < R; r++) {
                const int o = idx(base + r * Ns);
                re[o] = x[b * R + r].x;
                im[o] = x[b * R + r].y;
            }
        }
    }
}

template <class P, int s, typename T, class Idx>
__device__ __forceinline__ void lds_get(C2<T>* x, int tid, const T* re, const T* im, const Idx& idx) {
    constexpr int R = P::radix(s), NB = P::nb(s);
#pragma unroll
    for (int b = 0; b < NB; b++) {
        const int j = tid + b * P::TPL;
        if (P::full(s) || j < P::nbf(s)) {
#pragma unroll
            for (int r = 0; r < R; r++) {
                const int o = idx(j + r * (P::N / R));
                x[b * R + r] = mk<T>(re[o], im[o]);
            }
        }
    }
}

// a complex store; NT: with the nontemporal hint (one 16- or 8-byte vector
// store that does not allocate in the caches: the line is not read again
// before it would be evicted)
template <bool NT, typename T>
__device__ __forceinline__ void st_line(C2<T>* p, C2<T> v) {
    if constexpr (NT) {
        typedef T v2 __attribute__((ext_vector_type(2)));
        const v2 w = {v.x, v.y};
        __builtin_nontemporal_store(w, reinterpret_cast<v2*>(p));
    } else {
        *p = v;
    }
}

// Block barriers for the LDS exchanges.  SyncBlock is __syncthreads().
// SyncLds waits for this wave's LDS operations only and meets the other
// waves: global loads -- and loads into LDS (global_load_lds) for another
// buffer -- stay in flight across it, which a workgroup fence would drain.
// It orders LDS traffic only, so a kernel that uses it manages the waits for
// its loads into LDS itself.
struct SyncBlock {
    __device__ static __forceinline__ void run() { __syncthreads(); }
};
struct SyncLds {
    __device__ static __forceinline__ void run() { asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory"); }
};

// Whole-line transform.  On entry x holds the stage-0 inputs (see for_in),
// on exit the outputs (see for_out).  If the LDS arrays were used before the
// call, the caller must synchronise first.
template <class P, int SIGN, typename T, class Idx, class Sync = SyncBlock>
__device__ __forceinline__ void line_fft(C2<T>* x, int tid, T* re, T* im, const Idx& idx,
                                         const C2<T>* __restrict__ tw) {
    stage_bfly<P, 0, SIGN, T>(x, tid, tw);
    if constexpr (P::S > 1) {
        lds_put<P, 0, T>(x, tid, re, im, idx);
        Sync::run();
        lds_get<P, 1, T>(x, tid, re, im, idx);
        stage_bfly<P, 1, SIGN, T>(x, tid, tw);
    }
    if constexpr (P::S > 2) {
        Sync::run();
        lds_put<P, 1, T>(x, tid, re, im, idx);
        Sync::run();
        lds_get<P, 2, T>(x, tid, re, im, idx);
        stage_bfly<P, 2, SIGN, T>(x, tid, tw);
    }
    if constexpr (P::S > 3) {
        Sync::run();
        lds_put<P, 2, T>(x, tid, re, im, idx);
        Sync::run();
        lds_get<P, 3, T>(x, tid, re, im, idx);
        stage_bfly<P, 3, SIGN, T>(x, tid, tw);
    }
    if constexpr (P::S > 4) {
        Sync::run();
        lds_put<P, 3, T>(x, tid, re, im, idx);
        Sync::run();
        lds_get<P, 4, T>(x, tid, re, im, idx);
        stage_bfly<P, 4, SIGN, T>(x, tid, tw);
    }
}

// The same transform exchanging interleaved complex values through one LDS
// array (one 16-byte access per value in FP64 instead of two 8-byte ones):
// fewer LDS instructions, and with consecutive lanes on consecutive values
// (idx column-fastest) the accesses stay free of bank conflicts.
template <class P, int s, typename T, class Idx>
__device__ __forceinline__ void lds_put_c(const C2<T>* x, int tid, C2<T>* buf, const Idx& idx) {
    constexpr int R = P::radix(s), Ns = P::ns(s), NB = P::nb(s);
#pragma unroll
    for (int b = 0; b < NB; b++) {
        const int j = tid + b * P::TPL;
        if (P::full(s) || j < P::nbf(s)) {
            const int base = (j / Ns) * Ns * R + (j % Ns);
#pragma unroll
            for (int r = 0; r < R; r++) buf[idx(base + r * Ns)] = x[b * R + r];
        }
    }
}

template <class P, int s, typename T, class Idx>
__device__ __forceinline__ void lds_get_c(C2<T>* x, int tid, const C2<T>* buf, const Idx& idx) {
    constexpr int R = P::radix(s), NB = P::nb(s);
#pragma unroll
    for (int b = 0; b < NB; b++) {
        const int j = tid + b * P::TPL;
        if (P::full(s) || j < P::nbf(s)) {
#pragma unroll
            for (int r = 0; r < R; r++) x[b * R + r] = buf[idx(j + r * (P::N / R))];
        }
    }
}

// XOR-swizzled index of a line of 16-byte values (a permutation of 0..N-1
// for N a multiple of 64): bits 0..3 of p flipped by bits 3..5, so that the
// puts and gets of the radix-4 and radix-2 stages of a line held by one wave
// (32 or 64 threads) and the reads of a whole stage touch distinct LDS banks
// in groups of 8 and of 16 lanes; positions r 64 apart keep their offsets
struct SwzIdx {
    __device__ __forceinline__ int operator()(int p) const {
        const int h = p >> 3;
        return p ^ ((h & 3) | (((h ^ (h >> 1)) & 1) << 2) | ((h & 4) << 1));
    }
};

// Stage-major twiddle table: for each stage s with Ns > 1 the factors
// w^(k r step_s), r = 1..R-1, k = 0..Ns-1, at sm_off(s) + (r - 1) Ns + k, so
// the threads of a stage read consecutive entries (no bank conflicts); the
// factors are those stage_bfly reads from the natural table
template <class P, int s>
__host__ __device__ constexpr int sm_off() {
    if constexpr (s <= 1)
        return 0;
    else
        return sm_off<P, s - 1>() + (P::ns(s - 1) > 1 ? (P::radix(s - 1) - 1) * P::ns(s - 1) : 0);
}
template <class P>
__host__ __device__ constexpr int sm_size() {
    return sm_off<P, P::S>();
}

// the stage-major table from the natural one (entries t0, t0 + stride, ...)
template <class P, int s, typename T>
__device__ __forceinline__ void sm_fill(C2<T>* dst, const C2<T>* __restrict__ tw, int t0, int stride) {
    if constexpr (s < P::S) {
        constexpr int R = P::radix(s), Ns = P::ns(s);
        if constexpr (Ns > 1) {
            constexpr int step = P::N / (Ns * R);
            for (int e = t0; e < (R - 1) * Ns; e += stride) {
                const int r = e / Ns + 1, k = e % Ns;
                dst[sm_off<P, s>() + e] = tw[k * r * step];
            }
        }
        sm_fill<P, s + 1, T>(dst, tw, t0, stride);
    }
}

template <class P, int SIGN, typename T, class Idx, class Sync = SyncBlock>
__device__ __forceinline__ void line_fft_c(C2<T>* x, int tid, C2<T>* buf, const Idx& idx,
                                           const C2<T>* __restrict__ tw) {
    stage_bfly<P, 0, SIGN, T>(x, tid, tw);
    if constexpr (P::S > 1) {
        lds_put_c<P, 0, T>(x, tid, buf, idx);
        Sync::run();
        lds_get_c<P, 1, T>(x, tid, buf, idx);
        stage_bfly<P, 1, SIGN, T>(x, tid, tw);
    }
    if constexpr (P::S > 2) {
        Sync::run();
        lds_put_c<P, 1, T>(x, tid, buf, idx);
        Sync::run();
        lds_get_c<P, 2, T>(x, tid, buf, idx);
        stage_bfly<P, 2, SIGN, T>(x, tid, tw);
    }
    if constexpr (P::S > 3) {
        Sync::run();
        lds_put_c<P, 2, T>(x, tid, buf, idx);
        Sync::run();
        lds_get_c<P, 3, T>(x, tid, buf, idx);
        stage_bfly<P, 3, SIGN, T>(x, tid, tw);
    }
    if constexpr (P::S > 4) {
        Sync::run();
        lds_put_c<P, 3, T>(x, tid, buf, idx);
        Sync::run();
        lds_get_c<P, 4, T>(x, tid, buf, idx);
        stage_bfly<P, 4, SIGN, T>(x, tid, tw);
    }
}

// Per-thread stage twiddles held in registers: a thread's factors depend on
// its index alone (k = j mod Ns), so a kernel that transforms many lines with
// the same threads can load them once instead of reading the table in every
// stage.  Layout: stages with Ns > 1 in order, butterfly b, r = 1..R-1.
template <class P, int s>
__host__ __device__ constexpr int rt_off() {
    if constexpr (s <= 0)
        return 0;
    else
        return rt_off<P, s - 1>() + (P::ns(s - 1) > 1 ? P::nb(s - 1) * (P::radix(s - 1) - 1) : 0);
}
template <class P>
__host__ __device__ constexpr int rt_size() {
    return rt_off<P, P::S>();
}

// this thread's factors from the natural table (the ones stage_bfly reads)
template <class P, int s, typename T>
__device__ __forceinline__ void rt_load(C2<T>* w, int tid, const C2<T>* __restrict__ tw) {
    if constexpr (s < P::S) {
        constexpr int R = P::radix(s), Ns = P::ns(s), NB = P::nb(s);
        if constexpr (Ns > 1) {
            constexpr int step = P::N / (Ns * R);
#pragma unroll
            for (int b = 0; b < NB; b++) {
                const int j = tid + b * P::TPL;
                const int k = (P::full(s) || j < P::nbf(s)) ? j % Ns : 0;
#pragma unroll
                for (int r = 1; r < R; r++) w[rt_off<P, s>() + b * (R - 1) + r - 1] = tw[k * r * step];
            }
        }
        rt_load<P, s + 1, T>(w, tid, tw);
    }
}

template <class P, int s, int SIGN, typename T>
__device__ __forceinline__ void stage_bfly_r(C2<T>* x, int tid, const C2<T>* w) {
    constexpr int R = P::radix(s), Ns = P::ns(s), NB = P::nb(s);
#pragma unroll
    for (int b = 0; b < NB; b++) {
        const int j = tid + b * P::TPL;
        if (P::full(s) || j < P::nbf(s)) {
            C2<T>* v = x + b * R;
            if constexpr (Ns > 1) {
#pragma unroll
                for (int r = 1; r < R; r++) {
                    C2<T> t = w[rt_off<P, s>() + b * (R - 1) + r - 1];
                    if constexpr (SIGN > 0) t.y = -t.y;
                    v[r] = cmul<T>(v[r], t);
                }
            }
            Dft<T, R, SIGN>::run(v);
        }
    }
}

// line_fft_c with the stage twiddles from registers (rt_load)
template <class P, int SIGN, typename T, class Idx, class Sync = SyncBlock>
__device__ __forceinline__ void line_fft_cr(C2<T>* x, int tid, C2<T>* buf, const Idx& idx, const C2<T>* w) {
    stage_bfly_r<P, 0, SIGN, T>(x, tid, w);
    if constexpr (P::S > 1) {
        lds_put_c<P, 0, T>(x, tid, buf, idx);
        Sync::run();
        lds_get_c<P, 1, T>(x, tid, buf, idx);
        stage_bfly_r<P, 1, SIGN, T>(x, tid, w);
    }
    if constexpr (P::S > 2) {
        Sync::run();
        lds_put_c<P, 1, T>(x, tid, buf, idx);
        Sync::run();
        lds_get_c<P, 2, T>(x, tid, buf, idx);
        stage_bfly_r<P, 2, SIGN, T>(x, tid, w);
    }
    if constexpr (P::S > 3) {
        Sync::run();
        lds_put_c<P, 2, T>(x, tid, buf, idx);
        Sync::run();
        lds_get_c<P, 3, T>(x, tid, buf, idx);
        stage_bfly_r<P, 3, SIGN, T>(x, tid, w);
    }
    if constexpr (P::S > 4) {
        Sync::run();
        lds_put_c<P, 3, T>(x, tid, buf, idx);
        Sync::run();
        lds_get_c<P, 4, T>(x, tid, buf, idx);
        stage_bfly_r<P, 4, SIGN, T>(x, tid, w);
    }
}

// f(slot, pos) for every stage-0 input this thread owns
template <class P, class F>
__device__ __forceinline__ void for_in(int tid, F&& f) {
    constexpr int R = P::radix(0), NB = P::nb(0);
#pragma unroll
    for (int b = 0; b < NB; b++) {
        const int j = tid + b * P::TPL;
        if (P::full(0) || j < P::nbf(0)) {
#pragma unroll
            for (int r = 0; r < R; r++) f(b * R + r, j + r * (P::N / R));
        }
    }
}

// f(slot, pos) for every final output this thread owns
template <class P, class F>
__device__ __forceinline__ void for_out(int tid, F&& f) {
    constexpr int s = P::S - 1;
    constexpr int R = P::radix(s), NB = P::nb(s);
#pragma unroll
    for (int b = 0; b < NB; b++) {
        const int j = tid + b * P::TPL;
        if (P::full(s) || j < P::nbf(s)) {
#pragma unroll
            for (int r = 0; r < R; r++) f(b * R + r, j + r * (P::N / R));
        }
    }
}

// Plans used per length.  TPL = N / max radix; the tuple is (plan, lines per
// block in the contiguous pass, columns per block in the strided pass).
using P8 = FftPlan<8, 1, 8>;
using P12 = FftPlan<12, 3, 4, 3>;
using P16 = FftPlan<16, 4, 4, 4>;
using P24 = FftPlan<24, 3, 8, 3>;
using P32 = FftPlan<32, 4, 8, 4>;
using P48 = FftPlan<48, 12, 4, 4, 3>;
using P64 = FftPlan<64, 8, 8, 8>;
using P96 = FftPlan<96, 12, 8, 4, 3>;
using P128 = FftPlan<128, 8, 16, 8>;
using P256 = FftPlan<256, 16, 16, 16>;
using P512 = FftPlan<512, 64, 8, 8, 8>;
// lengths with the odd radices 5 and 7, and 1024
using P10 = FftPlan<10, 2, 5, 2>;
using P14 = FftPlan<14, 2, 7, 2>;
using P20 = FftPlan<20, 4, 5, 4>;
using P28 = FftPlan<28, 4, 7, 4>;
using P40 = FftPlan<40, 8, 5, 8>;
using P56 = FftPlan<56, 8, 7, 8>;
using P80 = FftPlan<80, 16, 5, 16>;
using P100 = FftPlan<100, 20, 5, 5, 4>;
using P1024 = FftPlan<1024, 64, 16, 4, 16>;

#define SH_FOR_EACH_FFT_SIZE(X) \
    X(8, P8, 64, 16)            \
    X(10, P10, 64, 32)          \
    X(12, P12, 64, 16)          \
    X(14, P14, 64, 32)          \
    X(16, P16, 64, 16)          \
    X(20, P20, 64, 16)          \
    X(24, P24, 64, 16)          \
    X(28, P28, 64, 16)          \
    X(32, P32, 64, 16)          \
    X(40, P40, 32, 16)          \
    X(48, P48, 16, 16)          \
    X(56, P56, 32, 16)          \
    X(64, P64, 32, 16)          \
    X(80, P80, 16, 16)          \
    X(96, P96, 16, 16)          \
    X(100, P100, 16, 16)        \
    X(128, P128, 32, 16)        \
    X(256, P256, 16, 16)        \
    X(512, P512, 4, 8)          \
    X(1024, P1024, 2, 2)

// Plans for the fused solver kernels: at most 9 complex values per thread
// (several fields are held at once), and first radix == last radix, so the
// output slot->position map of a transform equals the input map of the next
// one and a thread can chain an inverse transform, a pointwise product and a
// forward transform without an LDS reshuffle.
using Q12 = FftPlan<12, 4, 2, 3, 2>;
using Q24 = FftPlan<24, 8, 2, 3, 2, 2>;
using Q32 = FftPlan<32, 8, 4, 2, 4>;
using Q48 = FftPlan<48, 12, 4, 3, 4>;
using Q96 = FftPlan<96, 24, 4, 2, 3, 4>;
using Q128 = FftPlan<128, 16, 8, 2, 8>;
using Q256 = FftPlan<256, 32, 8, 4, 8>;

#define SH_FOR_EACH_FUSED_SIZE(X) \
    X(8, P8, 16)                  \
    X(12, Q12, 16)                \
    X(16, P16, 16)                \
    X(24, Q24, 16)                \
    X(32, Q32, 16)                \
    X(48, Q48, 16)                \
    X(64, P64, 16)                \
    X(96, Q96, 16)                \
    X(128, Q128, 16)              \
    X(256, Q256, 16)              \
    X(512, P512, 8)

// epilogue of a pass: optional scale and Nyquist zeroing (src/Scalar.cpp:336-370)
template <typename T>
struct Epi {
    T scale;
    int flags;   // 1: scale, 2: zero Nyquist
    int a_n;     // contiguous pass: line -> (a = line % a_n, b = line / a_n)
    int za, zb;  // Nyquist index for coordinates a and b (strided: a = column, b = outer)
};

// batched passes (fft3d.hip); dir = -1 forward, +1 backward
// role: ROLE_X/Y/Z (tuned variant of that role) or -1 (the default schedule)
template <typename T>
int pass_contig(C2<T>* data, int n, long nlines, long ldist, int count, long fstride, int dir,
                const Epi<T>& ep, hipStream_t st, int role);
template <typename T>
int pass_strided(C2<T>* data, int n, int ncols, long cdist, int nouter, long odist, long es,
                 int count, long fstride, int dir, const Epi<T>& ep, hipStream_t st, int role);
int pass_variant_count(int precision, int n, int role);
const char* pass_variant_name(int precision, int n, int role, int v);

// FP64 plans for the fused z kernel, which holds 7 fields per thread at
// once (3 velocity, 3 partial products, 1 in flight): at most 6 complex
// values per field keeps it under 128 VGPRs, so four waves share a SIMD.
using W64 = FftPlan<64, 16, 4, 4, 4>;
using W128 = FftPlan<128, 32, 4, 2, 4, 4>;
using W256 = FftPlan<256, 64, 4, 4, 4, 4>;
using W512 = FftPlan<512, 128, 4, 4, 2, 4, 4>;

#define SH_FOR_EACH_FUSED64_SIZE(X) \
    X(8, P8, 16)                    \
    X(12, Q12, 16)                  \
    X(16, P16, 16)                  \
    X(24, Q24, 16)                  \
    X(32, Q32, 16)                  \
    X(48, Q48, 16)                  \
    X(64, W64, 16)                  \
    X(96, Q96, 16)                  \
    X(128, W128, 16)                \
    X(256, W256, 16)                \
    X(512, W512, 8)

inline bool fft_size_supported(int n) {
    switch (n) {
#define SH_CASE(n_, P_, L_, C_) case n_:
        SH_FOR_EACH_FFT_SIZE(SH_CASE)
#undef SH_CASE
        return true;
        default:
            return false;
    }
}

// lengths with a compiled plan, a mixed-radix plan (fft_any.hip) or a
// Bluestein transform: every length the passes accept
bool fft_length_supported(int n);
// the pass over lengths without a compiled plan (fft_any.hip): line L of the
// batch starts at (L % a_n) a_d + (L / a_n) b_d, elements at stride es;
// strided selects the column-fastest lane order
template <typename T>
int pass_any(C2<T>* data, int n, long nlines, long a_n, long a_d, long b_d, long es, int count, long fstride,
             int dir, const Epi<T>& ep, bool strided, hipStream_t st);

inline bool fused_size_supported(int n) {
    switch (n) {
#define SH_CASE(n_, P_, C_) case n_:
        SH_FOR_EACH_FUSED_SIZE(SH_CASE)
#undef SH_CASE
        return true;
        default:
            return false;
    }
}

// ---------------------------------------------------------------- tunable variants
// Alternative radix/tile schedules for the lengths the workloads use.  Every
// variant computes the same transform; the autotuner (tune.hip) times them on
// the device and the dispatchers use its choice (index 0 = the default above).
using X512 = FftPlan<512, 32, 16, 2, 16>;

// strided pass: X(n, plan, columns per block)
// the same with nontemporal stores (X(n, plan, columns per block))
#define SH_STRIDED_NT_VARIANTS(X) X(256, Q256, 16) X(256, Q256, 8) X(256, W256, 16) X(512, X512, 16) X(512, P512, 8)
// with nontemporal loads: X(n, plan, columns per block, nontemporal stores)
// FP32 only (an FP64 tile of 32 columns of 512 would not fit in LDS):
// X(n, plan, columns per block, nontemporal stores, nontemporal loads)
#define SH_STRIDED_F32_VARIANTS(X) X(512, X512, 32, true, true) X(512, X512, 32, false, true) X(512, X512, 32, true, false)
#define SH_STRIDED_LNT_VARIANTS(X) \
    X(256, Q256, 16, true) X(256, Q256, 16, false) X(256, Q256, 8, false) X(512, X512, 16, true) X(512, X512, 16, false)
#define SH_STRIDED_VARIANTS(X) \
    X(64, P64, 16)             \
    X(64, P64, 32)             \
    X(64, W64, 16)             \
    X(96, P96, 16)             \
    X(96, P96, 32)             \
    X(96, Q96, 16)             \
    X(256, P256, 16)           \
    X(256, P256, 8)            \
    X(256, Q256, 16)           \
    X(256, Q256, 8)            \
    X(256, W256, 16)           \
    X(512, P512, 8)            \
    X(512, P512, 4)            \
    X(512, X512, 8)            \
    X(512, X512, 16)           \
    X(512, W512, 8)

// contiguous pass: X(n, plan, lines per block)
#define SH_CONTIG_VARIANTS(X) \
    X(64, P64, 32)            \
    X(64, P64, 16)            \
    X(64, W64, 16)            \
    X(96, P96, 16)            \
    X(96, Q96, 8)             \
    X(256, P256, 16)          \
    X(256, P256, 8)           \
    X(256, Q256, 8)           \
    X(256, W256, 4)           \
    X(512, P512, 4)           \
    X(512, P512, 2)           \
    X(512, X512, 8)           \
    X(512, X512, 4)           \
    X(512, W512, 2)

// contiguous pass with cache hints: X(n, plan, lines per block, nontemporal
// stores, nontemporal loads)
#define SH_CONTIG_HINT_VARIANTS(X)                                                                   \
    X(256, Q256, 8, true, false) X(256, Q256, 8, false, true) X(256, Q256, 8, true, true)             \
    X(512, P512, 2, true, false) X(512, P512, 2, false, true) X(512, P512, 2, true, true)

// pass roles, so y and z lines (different strides) are tuned separately;
// ROLE_CROSS_H is the fused z kernel of the real-field (half-spectrum) layout;
// ROLE_YCURL the solver's y inverse that forms the vorticity on load
enum { ROLE_X = 0, ROLE_Y = 1, ROLE_Z = 2, ROLE_CROSS = 3, ROLE_FOURIER = 4, ROLE_CROSS_H = 5, ROLE_YCURL = 6,
       NROLES = 7 };

// chosen variant for (precision, length, role); 0 when untuned
int tune_choice(int precision, int n, int role);
void tune_set(int precision, int n, int role, int variant, double ms);

}  // namespace sh
