// ns_step.hip -- the fractional step of exec/ns.cpp (AB2 on the nonlinear and
// pressure terms, Crank-Nicolson on the viscous term), two schedules:
//
//  * op-by-op (fused = 0, FP64): the reference's sequence of Scalar/Vector
//    operations (exec/ns.cpp:167-217), one kernel per operation, with the
//    temporaries the reference allocates per step kept resident instead.
//
//  * fused (fused = 1, FP64 or FP32): four kernels per step.
//      A  x-pass inverse of V and of omega = i k x V (curl formed on load);
//         runs once after a (re)start, later steps get it from E (every
//         step with a Fourier-update schedule that does not emit it)
//      B  y-pass inverse of the six fields
//      C  z-pass inverse of the six fields, complex pointwise u x omega,
//         z-pass forward of the three products -- all in registers
//      D  y-pass forward of the products
//      E  x-pass forward of the products + Nyquist zeroing + the whole
//         Fourier-space update (F, L, G = lap^-1 div(V+L+F), V - grad G,
//         CN divide) + |div V| partial sums, then A's work for the next
//         step from the new V still in registers
//    Per step this moves 48 field passes (768 N^3 bytes in FP64) against the
//    reference's 9 3D FFTs + ~20 array sweeps.  Every operation is mode-local
//    except the transforms, so the fusion changes no arithmetic beyond the
//    rounding of a different (equally exact) operation order.
//
// State (Fourier space): V1, N0 (= cross0), and six work fields W.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <type_traits>
#include <vector>

#include "fft_core.hpp"

namespace sh {

template <typename T>
int fft3d(void* data, int count, long fstride, int nx, int ny, int nz, int dir, int flags,
          void* stream);
int reduce_z(const void* f, long n, int mode, double* result, hipStream_t st);

// ---------------------------------------------------------------- fused kernels
template <typename T>
struct StepCoef {
    T Kh;          // K/2                   (exec/ns.cpp:171)
    T c1;          // nu*K/2                (exec/ns.cpp:174; nu*K in exec/ns2.cpp:178)
    double c1d;    // the same in FP64 for the viscous divide (exec/ns.cpp:186; ns2.cpp:187)
    int be;        // 1: exec/ns2.cpp's backward-Euler viscous term (V = Vhalf / (1 + nu K k^2))
    T inv_n3;      // 1/(Nx Ny Nz)          (src/Scalar.cpp:223-227)
    int dealias;   // 1: products with some |k_i| > N_i/3 are dropped (2/3 rule);
                   // 2: products scaled by prod_i exp(-36 (|k_i| / (N_i/2))^36)
};

// the smooth dealiasing factor of one axis, exp(-36 (|k| / (n/2))^36)
template <typename T>
__host__ __device__ __forceinline__ T expfilter(T k, int n) {
    const double r = fabs((double)k) / (0.5 * n);
    const double r4 = (r * r) * (r * r);
    const double r8 = r4 * r4, r16 = r8 * r8;
    return T(exp(-36.0 * (r16 * r16) * r4));  // r^32 r^4 = r^36
}

__device__ __forceinline__ double fmax_d(double a, double b) { return a > b ? a : b; }

template <int N>
struct LdsLine {
    static constexpr int LS = N + N / 16 + 1;
};

// Layout of an ensemble of B equal boxes (B = 1 for a single solver): field f
// of box b starts at (f*B + b)*Nx*Ny*Nz, so F below is the field stride B*N
// and box b is blockIdx.y, offset b*N.  The y passes then cover all boxes of a
// field group in one strided launch.

// A: inverse x-pass of V (3 fields) and omega = i k x V (3 fields) into W.
template <typename T, class P, int LPB>
__global__ __launch_bounds__(P::TPL* LPB * 6) void k_ns_inv_x(const C2<T>* __restrict__ V,
                                                              C2<T>* __restrict__ W,
                                                              const C2<T>* __restrict__ tw,
                                                              int ny, int nz, int nzp, long F, long boff) {
    V += blockIdx.y * boff;
    W += blockIdx.y * boff;
    constexpr int N = P::N;
    constexpr int LS = LdsLine<N>::LS;
    __shared__ T sre[6 * LPB * LS];
    __shared__ T sim[6 * LPB * LS];
    const int tid = threadIdx.x % P::TPL;
    const int ll = threadIdx.x / P::TPL;  // 0 .. 6*LPB-1
    const int g = ll / LPB;               // field group: 0-2 velocity, 3-5 vorticity
    const int l = ll % LPB;
    const long nlines = (long)ny * nzp;  // nzp = nz, or nz/2 in the half-spectrum layout
    const long line = (long)blockIdx.x * LPB + l;
    const bool valid = line < nlines;
    const long lc = valid ? line : 0;
    const T k1 = freq<T>((int)(lc % ny), ny);
    const T k2 = freq<T>((int)(lc / ny), nz);
    const C2<T>* vb = V + lc * N;
    C2<T> x[P::E];
    if (g < 3) {
        for_in<P>(tid, [&](int s, int pos) { x[s] = vb[g * F + pos]; });
    } else if (g == 3) {  // i k1 V2 - i k2 V1
        for_in<P>(tid, [&](int s, int pos) { x[s] = csub<T>(imul<T>(k1, vb[2 * F + pos]), imul<T>(k2, vb[F + pos])); });
    } else if (g == 4) {  // i k2 V0 - i k0 V2
        for_in<P>(tid, [&](int s, int pos) {
            const T k0 = freq<T>(pos, N);
            x[s] = csub<T>(imul<T>(k2, vb[pos]), imul<T>(k0, vb[2 * F + pos]));
        });
    } else {  // i k0 V1 - i k1 V0
        for_in<P>(tid, [&](int s, int pos) {
            const T k0 = freq<T>(pos, N);
            x[s] = csub<T>(imul<T>(k0, vb[F + pos]), imul<T>(k1, vb[pos]));
        });
    }
    auto idx = [&](int pos) { return ll * LS + pos + (pos >> 4); };
    line_fft<P, 1, T>(x, tid, sre, sim, idx, tw);
    if (!valid) return;
    C2<T>* wb = W + g * F + lc * N;
    for_out<P>(tid, [&](int s, int pos) { wb[pos] = x[s]; });
}

// C: z-pass inverse of W[0..6), u x omega, z-pass forward into W[0..3).
// (W may be offset to a first column; ncx columns of rows nx long are processed)
template <typename T, class P, int LPB>
__global__ __launch_bounds__(P::TPL* LPB) void k_ns_cross_z(C2<T>* __restrict__ W,
                                                            const C2<T>* __restrict__ tw, int nx, int ncx,
                                                            int ny, long F, T scale, long boff) {
    static_assert(P::radix(0) == P::radix(P::S - 1) && P::nb(0) == P::nb(P::S - 1),
                  "fused z kernel needs a symmetric plan");
    W += blockIdx.y * boff;
    constexpr int CS = LPB + 1;
    constexpr int E = P::E;
    __shared__ T sre[P::N * CS];
    __shared__ T sim[P::N * CS];
    const int c = threadIdx.x % LPB;
    const int tid = threadIdx.x / LPB;
    const int tiles = (ncx + LPB - 1) / LPB;
    const int j = blockIdx.x / tiles;
    const int i = (blockIdx.x % tiles) * LPB + c;
    const bool valid = i < ncx && j < ny;
    // 32-bit element offsets from the uniform base W (6 fields of at most
    // 2^31/6 elements; checked by the launcher) halve the address registers
    const unsigned es = (unsigned)nx * (unsigned)ny;
    const unsigned uF = (unsigned)F;
    const unsigned cb = valid ? (unsigned)i + (unsigned)j * (unsigned)nx : 0u;
    auto idx = [&](int pos) { return pos * CS + c; };
    auto load = [&](int f, C2<T>* x) {
        for_in<P>(tid, [&](int s, int pos) { x[s] = W[cb + (unsigned)f * uF + (unsigned)pos * es]; });
    };
    // inverse z transform of x (scaled by 1/(Nx Ny Nz)); the loads of the
    // field after it were issued beforehand so they overlap the LDS phases
    auto inv = [&](C2<T>* x, bool first) {
        if (!first) __syncthreads();
        line_fft<P, 1, T>(x, tid, sre, sim, idx, tw);
#pragma unroll
        for (int s = 0; s < E; s++) x[s] = rmul<T>(scale, x[s]);
    };
    // ret0 = u1 w2 - u2 w1, ret1 = u2 w0 - u0 w2, ret2 = u0 w1 - u1 w0 (src/Vector.cpp:79-84).
    // Fields arrive in the order w0 w1 u2 w2 u0 u1 so at most 6 fields are
    // live; each a - b is accumulated as a + (-b) or (-b) + a, which round
    // identically to a - b.
    C2<T> w[3][E], o[3][E], x[E], pf[E];
    load(3, w[0]);
    load(4, w[1]);
    inv(w[0], true);
    load(2, pf);
    inv(w[1], false);
    load(5, w[2]);
    inv(pf, false);  // u2
#pragma unroll
    for (int s = 0; s < E; s++) {
        o[0][s] = csub<T>(mk<T>(T(0), T(0)), cmul<T>(pf[s], w[1][s]));
        o[1][s] = cmul<T>(pf[s], w[0][s]);
    }
    load(0, x);
    inv(w[2], false);
    load(1, pf);
    inv(x, false);  // u0
#pragma unroll
    for (int s = 0; s < E; s++) {
        o[1][s] = csub<T>(o[1][s], cmul<T>(x[s], w[2][s]));
        o[2][s] = cmul<T>(x[s], w[1][s]);
    }
    inv(pf, false);  // u1
#pragma unroll
    for (int s = 0; s < E; s++) {
        o[0][s] = cadd<T>(o[0][s], cmul<T>(pf[s], w[2][s]));
        o[2][s] = csub<T>(o[2][s], cmul<T>(pf[s], w[0][s]));
    }
#pragma unroll
    for (int d = 0; d < 3; d++) {
#pragma unroll
        for (int s = 0; s < E; s++) x[s] = o[d][s];
        __syncthreads();
        line_fft<P, -1, T>(x, tid, sre, sim, idx, tw);
        if (valid) for_out<P>(tid, [&](int s, int pos) { W[cb + (unsigned)d * uF + (unsigned)pos * es] = x[s]; });
    }
}

// C in the half-spectrum layout (real physical fields).  A z-line holds
// X[k], k = 0..M-1, M = nz/2: X[M] is the Nyquist plane forward() zeroes and
// X[nz-k] = conj X[k].  The length-nz real transforms run as length-M complex
// ones (P is a length-M plan) on packed pairs of grid points:
//   inverse: Z[k] = (X[k] + conj X[M-k]) + i e^{2 pi i k/nz} (X[k] - conj X[M-k]);
//            the length-M inverse of Z is y[2m] + i y[2m+1];
//   forward: Z = the length-M forward of c[2m] + i c[2m+1];
//            X[k] = ((Z[k] + conj Z[M-k]) - i e^{-2 pi i k/nz} (Z[k] - conj Z[M-k])) / 2.
// The cross product acts on the real and the imaginary parts separately (two
// grid points); the imaginary parts the reference's complex product carries
// are the rounding noise of real fields and are not formed.
template <typename T, class P, int LPB, bool NT = false>
__global__ __launch_bounds__(P::TPL* LPB) void k_ns_cross_zh(C2<T>* __restrict__ W,
                                                             const C2<T>* __restrict__ tw,
                                                             const C2<T>* __restrict__ twn, int nx,
                                                             int ncx, int ny, long F, T scale, long boff) {
    static_assert(P::radix(0) == P::radix(P::S - 1) && P::nb(0) == P::nb(P::S - 1),
                  "fused z kernel needs a symmetric plan");
    W += blockIdx.y * boff;
    constexpr int M = P::N;
    constexpr int CS = LPB + 1;
    constexpr int E = P::E;
    __shared__ T sre[M * CS];
    __shared__ T sim[M * CS];
    const int c = threadIdx.x % LPB;
    const int tid = threadIdx.x / LPB;
    const int tiles = (ncx + LPB - 1) / LPB;
    const int j = blockIdx.x / tiles;
    const int i = (blockIdx.x % tiles) * LPB + c;
    const bool valid = i < ncx && j < ny;
    const unsigned es = (unsigned)nx * (unsigned)ny;
    const unsigned uF = (unsigned)F;
    const unsigned cb = valid ? (unsigned)i + (unsigned)j * (unsigned)nx : 0u;
    auto idx = [&](int pos) { return pos * CS + c; };
    auto load = [&](int f, C2<T>* x) {
        for_in<P>(tid, [&](int s, int pos) { x[s] = W[cb + (unsigned)f * uF + (unsigned)pos * es]; });
    };
    // X -> Z (X[M-k] is held by another thread: one LDS round trip), the
    // length-M inverse, and the 1/(Nx Ny Nz) scale
    auto inv = [&](C2<T>* x) {
        __syncthreads();
        for_in<P>(tid, [&](int s, int pos) {
            sre[idx(pos)] = x[s].x;
            sim[idx(pos)] = x[s].y;
        });
        __syncthreads();
        for_in<P>(tid, [&](int s, int pos) {
            const C2<T> a = x[s];
            const C2<T> b = pos == 0 ? mk<T>(T(0), T(0)) : mk<T>(sre[idx(M - pos)], sim[idx(M - pos)]);
            const C2<T> sum = mk<T>(a.x + b.x, a.y - b.y);  // X[k] + conj X[M-k]
            const C2<T> dif = mk<T>(a.x - b.x, a.y + b.y);  // X[k] - conj X[M-k]
            const C2<T> w = twn[pos];                       // e^{-2 pi i k/nz}
            const T tx = w.x * dif.x + w.y * dif.y;         // t = conj(w) dif
            const T ty = w.x * dif.y - w.y * dif.x;
            x[s] = mk<T>(sum.x - ty, sum.y + tx);           // Z = sum + i t
        });
        __syncthreads();
        line_fft<P, 1, T>(x, tid, sre, sim, idx, tw);
#pragma unroll
        for (int s = 0; s < E; s++) x[s] = rmul<T>(scale, x[s]);
    };
    // elementwise product of two packed pairs of real values
    auto pmul = [](C2<T> a, C2<T> b) { return mk<T>(a.x * b.x, a.y * b.y); };
    // r0 = u1 w2 - u2 w1, r1 = u2 w0 - u0 w2, r2 = u0 w1 - u1 w0 (src/Vector.cpp:79-84),
    // fields in the order w0 w1 u2 w2 u0 u1 (at most 6 live), as kernel C
    C2<T> w[3][E], o[3][E], x[E], pf[E];
    load(3, w[0]);
    load(4, w[1]);
    inv(w[0]);
    load(2, pf);
    inv(w[1]);
    load(5, w[2]);
    inv(pf);  // u2
#pragma unroll
    for (int s = 0; s < E; s++) {
        o[0][s] = csub<T>(mk<T>(T(0), T(0)), pmul(pf[s], w[1][s]));
        o[1][s] = pmul(pf[s], w[0][s]);
    }
    load(0, x);
    inv(w[2]);
    load(1, pf);
    inv(x);  // u0
#pragma unroll
    for (int s = 0; s < E; s++) {
        o[1][s] = csub<T>(o[1][s], pmul(x[s], w[2][s]));
        o[2][s] = pmul(x[s], w[1][s]);
    }
    inv(pf);  // u1
#pragma unroll
    for (int s = 0; s < E; s++) {
        o[0][s] = cadd<T>(o[0][s], pmul(pf[s], w[2][s]));
        o[2][s] = csub<T>(o[2][s], pmul(pf[s], w[0][s]));
    }
#pragma unroll
    for (int d = 0; d < 3; d++) {
#pragma unroll
        for (int s = 0; s < E; s++) x[s] = o[d][s];
        __syncthreads();
        line_fft<P, -1, T>(x, tid, sre, sim, idx, tw);
        // Z -> X (Z[M-k] from another thread through LDS)
        __syncthreads();
        for_out<P>(tid, [&](int s, int pos) {
            sre[idx(pos)] = x[s].x;
            sim[idx(pos)] = x[s].y;
        });
        __syncthreads();
        if (valid)
            for_out<P>(tid, [&](int s, int pos) {
                const C2<T> a = x[s];
                const int mp = pos == 0 ? 0 : M - pos;
                const C2<T> b = mk<T>(sre[idx(mp)], sim[idx(mp)]);
                const C2<T> sum = mk<T>(a.x + b.x, a.y - b.y);  // Z[k] + conj Z[M-k]
                const C2<T> dif = mk<T>(a.x - b.x, a.y + b.y);  // Z[k] - conj Z[M-k]
                const C2<T> w = twn[pos];
                const T tx = w.x * dif.x - w.y * dif.y;  // t = w dif
                const T ty = w.x * dif.y + w.y * dif.x;
                st_line<NT, T>(W + (cb + (unsigned)d * uF + (unsigned)pos * es),
                               mk<T>(T(0.5) * (sum.x + ty), T(0.5) * (sum.y - tx)));
            });
    }
}

// ---- cross-lane exchanges
// A line's values spread over a wave's lanes and registers, lane l and
// register r holding position l + 64 r (or, for a line over 32 lanes, l + 32
// r): a radix-4 stage's data movement is then a swap of two lane-index bits
// with the two register-index bits, done with cross-lane instructions (lane
// bits 5, 4: v_permlane32_swap / v_permlane16_swap; bits 3, 2: DPP row shifts
// by 8 and 4 writing only the lanes whose bit is set or clear, by bank mask;
// bits 1, 0: DPP quad permutations and a select), with no LDS traffic and no
// barrier.
//
// xw_swap1<B>: a and b are one dword of the registers whose register-index bit
// is 0 and 1; afterwards a holds, on lanes with bit B set, b of the lane 2^B
// below, and b, on lanes with bit B clear, a of the lane 2^B above.
template <int B>
__device__ __forceinline__ void xw_swap1(unsigned& a, unsigned& b, int lane) {
    const unsigned ua = a, ub = b;
    if constexpr (B == 5) {
        const auto r = __builtin_amdgcn_permlane32_swap(ua, ub, false, false);
        a = r[0];
        b = r[1];
    } else if constexpr (B == 4) {
        const auto r = __builtin_amdgcn_permlane16_swap(ua, ub, false, false);
        a = r[0];
        b = r[1];
    } else if constexpr (B == 3) {
        a = __builtin_amdgcn_update_dpp(ua, ub, 0x118, 0xF, 0xC, false);  // row_shr:8, lanes 8-15 of a row
        b = __builtin_amdgcn_update_dpp(ub, ua, 0x108, 0xF, 0x3, false);  // row_shl:8, lanes 0-7
    } else if constexpr (B == 2) {
        a = __builtin_amdgcn_update_dpp(ua, ub, 0x114, 0xF, 0xA, false);  // row_shr:4, lanes 4-7, 12-15
        b = __builtin_amdgcn_update_dpp(ub, ua, 0x104, 0xF, 0x5, false);  // row_shl:4, lanes 0-3, 8-11
    } else {
        constexpr int QP = B == 1 ? 0x4E : 0xB1;  // quad_perm [2,3,0,1] (lane ^ 2), [1,0,3,2] (lane ^ 1)
        const unsigned pa = __builtin_amdgcn_mov_dpp(ua, QP, 0xF, 0xF, false);
        const unsigned pb = __builtin_amdgcn_mov_dpp(ub, QP, 0xF, 0xF, false);
        const bool hi = (lane >> B) & 1;
        a = hi ? pb : ua;
        b = hi ? ub : pa;
    }
}

// the same for one real part or imaginary part (bit-exact moves)
template <int B>
__device__ __forceinline__ void xw_swapv(float& a, float& b, int lane) {
    unsigned ua = __float_as_uint(a), ub = __float_as_uint(b);
    xw_swap1<B>(ua, ub, lane);
    a = __uint_as_float(ua);
    b = __uint_as_float(ub);
}
template <int B>
__device__ __forceinline__ void xw_swapv(double& a, double& b, int lane) {
    unsigned al = (unsigned)__double2loint(a), ah = (unsigned)__double2hiint(a);
    unsigned bl = (unsigned)__double2loint(b), bh = (unsigned)__double2hiint(b);
    xw_swap1<B>(al, bl, lane);
    xw_swap1<B>(ah, bh, lane);
    a = __hiloint2double((int)ah, (int)al);
    b = __hiloint2double((int)bh, (int)bl);
}

// lane bits (H, L) <-> register-index bits (1, 0) of x[0..3]: afterwards
// lane bits (H, L) hold what the register index was, and the register index
// what lane bits (H, L) were
template <int H, int L, typename T>
__device__ __forceinline__ void xw_swap2(C2<T>* x, int lane) {
#pragma unroll
    for (int p = 0; p < 2; p++) {
        xw_swapv<H>(x[p].x, x[p + 2].x, lane);
        xw_swapv<H>(x[p].y, x[p + 2].y, lane);
    }
#pragma unroll
    for (int p = 0; p < 4; p += 2) {
        xw_swapv<L>(x[p].x, x[p + 1].x, lane);
        xw_swapv<L>(x[p].y, x[p + 1].y, lane);
    }
}

// ---- 128-point transforms of kernel C with one LDS exchange
// A block of four waves transforms eight z lines (one per column c = lane % 8)
// of 128 points; thread (wave w, lane l) is line position t = 8 w + l / 8 and
// register r holds position t + 32 r, the layout the slot reads and the
// product stores use.  Of the transform's four stages (radix 4, 4, 4, 2 in
// decimation in frequency), only the first moves data between waves: one
// exchange through LDS (a transpose of the wave index with the register
// index, contiguous 1 KiB per wave instruction) and two barriers.  The
// second stage's data movement is the swap of register bits with lane bits 5,
// 4 (t bits 2, 1), the third's that of register bit 0 with lane bit 3 (t bit
// 0).
//   zw_inv: natural input (X[t + 32 r]), sign +1, output x[n] at
//     n = w + 4 ((l >> 4) & 3) + 16 ((l >> 3) & 1) + 32 (r >> 1) + 64 (r & 1);
//   zw_fwd: the inverse process (input in that order, sign -1, natural output,
//     unnormalised): pointwise products of zw_inv outputs need no reordering.
// Twiddles (e^{-2 pi i k / 128}, conjugated for zw_inv) in registers: a[j] at
// k = j t, b[j] at k = 4 j (t % 8), c[j] at k = 16 j (t % 2).
template <typename T>
struct ZwTw {
    C2<T> a[3], b[3], c[3];
    __device__ __forceinline__ void load(const C2<T>* __restrict__ twg, int t) {
#pragma unroll
        for (int j = 1; j < 4; j++) {
            a[j - 1] = twg[(j * t) & 127];
            b[j - 1] = twg[(4 * j * (t & 7)) & 127];
            c[j - 1] = twg[(16 * j * (t & 1)) & 127];
        }
    }
};

// x[1..3] times w[0..2] (CONJ: times their conjugates)
template <bool CONJ, typename T>
__device__ __forceinline__ void zw_twiddle(C2<T>* x, const C2<T>* w) {
#pragma unroll
    for (int r = 1; r < 4; r++) {
        const C2<T> a = x[r], b = w[r - 1];
        x[r] = CONJ ? mk<T>(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y) : cmul<T>(a, b);
    }
}

// the exchange: register r of (wave w, lane l) to entry 256 r + 64 w + l, then
// register r from entry 256 w + 64 r + l (the caller has synchronised the
// previous users of sx)
template <typename T>
__device__ __forceinline__ void zw_exchange(C2<T>* x, C2<T>* sx, int wave, int lane) {
#pragma unroll
    for (int r = 0; r < 4; r++) sx[256 * r + 64 * wave + lane] = x[r];
    SyncLds::run();
#pragma unroll
    for (int r = 0; r < 4; r++) x[r] = sx[256 * wave + 64 * r + lane];
}

// register bit 0 <-> lane bit 3, then the radix-2 butterflies over register bit 0
template <typename T>
__device__ __forceinline__ void zw_swap3(C2<T>* x, int lane) {
#pragma unroll
    for (int p = 0; p < 4; p += 2) {
        xw_swapv<3>(x[p].x, x[p + 1].x, lane);
        xw_swapv<3>(x[p].y, x[p + 1].y, lane);
    }
}
template <typename T>
__device__ __forceinline__ void zw_radix2(C2<T>* x) {
#pragma unroll
    for (int p = 0; p < 4; p += 2) Dft<T, 2, 1>::run(x + p);
}

template <typename T>
__device__ __forceinline__ void zw_inv(C2<T>* x, const ZwTw<T>& t, C2<T>* sx, int wave, int lane) {
    Dft<T, 4, 1>::run(x);
    zw_twiddle<true, T>(x, t.a);
    zw_exchange<T>(x, sx, wave, lane);
    Dft<T, 4, 1>::run(x);
    zw_twiddle<true, T>(x, t.b);
    xw_swap2<5, 4, T>(x, lane);
    Dft<T, 4, 1>::run(x);
    zw_twiddle<true, T>(x, t.c);
    zw_swap3<T>(x, lane);
    zw_radix2<T>(x);
}

template <typename T>
__device__ __forceinline__ void zw_fwd(C2<T>* x, const ZwTw<T>& t, C2<T>* sx, int wave, int lane) {
    zw_radix2<T>(x);
    zw_swap3<T>(x, lane);
    zw_twiddle<false, T>(x, t.c);
    Dft<T, 4, -1>::run(x);
    xw_swap2<5, 4, T>(x, lane);
    zw_twiddle<false, T>(x, t.b);
    Dft<T, 4, -1>::run(x);
    zw_exchange<T>(x, sx, wave, lane);
    zw_twiddle<false, T>(x, t.a);
    Dft<T, 4, -1>::run(x);
}

// The same with eight values per thread, for sixteen 128-point lines (one per
// column c = lane % 16) on four waves: thread (wave w, lane l) is line position
// t = 4 w + l / 16 and register r holds position t + 16 r.  Radix 8 over the
// registers; one LDS exchange of register bits 1, 0 with the wave index
// (register bit 2 stays); radix 4 over register bits 1, 0 (positions 4 w);
// the swap of register bits 1, 0 with lane bits 5, 4 (positions l / 16);
// radix 4.
//   zw8_inv: natural input, sign +1, output x[n] at
//     n = w + 4 (r >> 2) + 8 ((l >> 4) & 3) + 32 (r & 3);
//   zw8_fwd: the inverse process (that order in, sign -1, natural out).
// Twiddles (e^{-2 pi i k / 128}, conjugated for zw8_inv): a[j - 1] at k = j t
// (j = 1..7), b[j - 1] at k = 8 j (t % 4) (j = 1..3).
template <typename T>
struct ZwTw8 {
    C2<T> a[7], b[3];
    __device__ __forceinline__ void load(const C2<T>* __restrict__ twg, int t) {
#pragma unroll
        for (int j = 1; j < 8; j++) a[j - 1] = twg[(j * t) & 127];
#pragma unroll
        for (int j = 1; j < 4; j++) b[j - 1] = twg[(8 * j * (t & 3)) & 127];
    }
};

// register r = 4 r2 + q of wave w to entry 256 r + 64 w + l, then register
// 4 r2 + q of wave w from entry 256 (4 r2 + w) + 64 q + l
template <typename T>
__device__ __forceinline__ void zw8_exchange(C2<T>* x, C2<T>* sx, int wave, int lane) {
#pragma unroll
    for (int r = 0; r < 8; r++) sx[256 * r + 64 * wave + lane] = x[r];
    SyncLds::run();
#pragma unroll
    for (int r = 0; r < 8; r++) x[r] = sx[256 * (4 * (r >> 2) + wave) + 64 * (r & 3) + lane];
}

template <typename T>
__device__ __forceinline__ void zw8_inv(C2<T>* x, const ZwTw8<T>& t, C2<T>* sx, int wave, int lane) {
    Dft<T, 8, 1>::run(x);
#pragma unroll
    for (int r = 1; r < 8; r++) {
        const C2<T> a = x[r], b = t.a[r - 1];
        x[r] = mk<T>(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
    }
    zw8_exchange<T>(x, sx, wave, lane);
    Dft<T, 4, 1>::run(x);
    Dft<T, 4, 1>::run(x + 4);
    zw_twiddle<true, T>(x, t.b);
    zw_twiddle<true, T>(x + 4, t.b);
    xw_swap2<5, 4, T>(x, lane);
    xw_swap2<5, 4, T>(x + 4, lane);
    Dft<T, 4, 1>::run(x);
    Dft<T, 4, 1>::run(x + 4);
}

template <typename T>
__device__ __forceinline__ void zw8_fwd(C2<T>* x, const ZwTw8<T>& t, C2<T>* sx, int wave, int lane) {
    Dft<T, 4, -1>::run(x);
    Dft<T, 4, -1>::run(x + 4);
    xw_swap2<5, 4, T>(x, lane);
    xw_swap2<5, 4, T>(x + 4, lane);
    zw_twiddle<false, T>(x, t.b);
    zw_twiddle<false, T>(x + 4, t.b);
    Dft<T, 4, -1>::run(x);
    Dft<T, 4, -1>::run(x + 4);
    zw8_exchange<T>(x, sx, wave, lane);
#pragma unroll
    for (int r = 1; r < 8; r++) x[r] = cmul<T>(x[r], t.a[r - 1]);
    Dft<T, 8, -1>::run(x);
}

// The same for sixteen 256-point lines (one per column c = lane % 16) on
// eight waves: thread (wave w, lane l) is line position t = 4 w + l / 16 and
// register r holds position t + 32 r.  Radix 8 over the registers; one LDS
// exchange transposing the register index with the wave index; radix 8; the
// swap of register bits 1, 0 with lane bits 5, 4; radix 4.
//   zs_inv: natural input, sign +1, output x[n] at
//     n = w + 8 (4 (r >> 2) + ((l >> 4) & 3)) + 64 (r & 3);
//   zs_fwd: the inverse process (that order in, sign -1, natural out).
// Twiddles (e^{-2 pi i k / 256}, conjugated for zs_inv): a[j - 1] at k = j t,
// b[j - 1] at k = 8 j (t % 4) (j = 1..7).
template <typename T>
struct ZsTw {
    C2<T> a[7], b[7];
    __device__ __forceinline__ void load(const C2<T>* __restrict__ twg, int t) {
#pragma unroll
        for (int j = 1; j < 8; j++) {
            a[j - 1] = twg[(j * t) & 255];
            b[j - 1] = twg[(8 * j * (t & 3)) & 255];
        }
    }
};

// x[1..7] times w[0..6] (CONJ: times their conjugates)
template <bool CONJ, typename T>
__device__ __forceinline__ void zs_twiddle(C2<T>* x, const C2<T>* w) {
#pragma unroll
    for (int r = 1; r < 8; r++) {
        const C2<T> a = x[r], b = w[r - 1];
        x[r] = CONJ ? mk<T>(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y) : cmul<T>(a, b);
    }
}

// register r of wave w to entry 512 r + 64 w + l, then register r from entry
// 512 w + 64 r + l
template <typename T>
__device__ __forceinline__ void zs_exchange(C2<T>* x, C2<T>* sx, int wave, int lane) {
#pragma unroll
    for (int r = 0; r < 8; r++) sx[512 * r + 64 * wave + lane] = x[r];
    SyncLds::run();
#pragma unroll
    for (int r = 0; r < 8; r++) x[r] = sx[512 * wave + 64 * r + lane];
}

template <typename T>
__device__ __forceinline__ void zs_inv(C2<T>* x, const ZsTw<T>& t, C2<T>* sx, int wave, int lane) {
    Dft<T, 8, 1>::run(x);
    zs_twiddle<true, T>(x, t.a);
    zs_exchange<T>(x, sx, wave, lane);
    Dft<T, 8, 1>::run(x);
    zs_twiddle<true, T>(x, t.b);
    xw_swap2<5, 4, T>(x, lane);
    xw_swap2<5, 4, T>(x + 4, lane);
    Dft<T, 4, 1>::run(x);
    Dft<T, 4, 1>::run(x + 4);
}

template <typename T>
__device__ __forceinline__ void zs_fwd(C2<T>* x, const ZsTw<T>& t, C2<T>* sx, int wave, int lane) {
    Dft<T, 4, -1>::run(x);
    Dft<T, 4, -1>::run(x + 4);
    xw_swap2<5, 4, T>(x, lane);
    xw_swap2<5, 4, T>(x + 4, lane);
    zs_twiddle<false, T>(x, t.b);
    Dft<T, 8, -1>::run(x);
    zs_exchange<T>(x, sx, wave, lane);
    zs_twiddle<false, T>(x, t.a);
    Dft<T, 8, -1>::run(x);
}

// C of the half-spectrum layout with its inputs streamed into LDS (FP64).
// The same arithmetic as k_ns_cross_zh; what changes is how the six input
// fields arrive.  Each block loops over column tiles (a persistent grid of
// resident blocks) and keeps three field slots in LDS that gfx950's direct
// global->LDS loads (global_load_lds_dwordx4, no registers) fill two fields
// ahead: while a tile's field is transformed, the next two are in flight, and
// during a tile's three forward transforms the next tile's first three are.
// The FFT exchanges synchronise with SyncLds, which does not drain those
// loads, and each slot is read only after an explicit wait for its own loads
// (vmcnt counts this wave's loads into LDS and its stores, in issue order)
// and a barrier.  The half-spectrum packing reads X[k] and X[M-k] straight
// from the slot, so it needs no exchange of its own.  The twiddle tables are
// copied into LDS once per block, so the only vector-memory operations left
// are the slot loads and the product stores.  HALF = false: the same kernel
// for the full spectrum (the plain C2C method, k_ns_cross_z's arithmetic): a
// z-line is one length-nz transform, with no packing, the products are full
// complex products, and the forward transforms are stored as they are.
template <int N>
__device__ __forceinline__ void vm_wait_barrier() {
    asm volatile("s_waitcnt vmcnt(%0)\n\ts_barrier" ::"n"(N) : "memory");
}

template <typename T>
__device__ __forceinline__ unsigned lds_offset(const T* p) {
    return (unsigned)(uintptr_t)(const __attribute__((address_space(3))) T*)p;
}

template <class P, int LPB, bool NT = false, bool IL = false, bool HALF = true, typename T = double, int RT = 0,
          int LAUX = 0>
__global__ __launch_bounds__(P::TPL* LPB)
__attribute__((amdgpu_waves_per_eu(RT >= 4 ? 2 : sizeof(T) == 4 ? 4 : 1)))
void k_ns_cross_zd(C2<T>* __restrict__ W,
                                                             const C2<T>* __restrict__ twg,
                                                             const C2<T>* __restrict__ twng, int nx,
                                                             int ncx, int ny, long F, T scale, long boff,
                                                             int ntile) {
    constexpr int M = P::N;
    constexpr int R0 = P::radix(0);
    constexpr int Q = M / R0;
    constexpr int CS = LPB + 1;
    constexpr int E = P::E;
    constexpr int NTH = P::TPL * LPB;
    constexpr int NW = NTH / 64;
    constexpr unsigned VB = sizeof(C2<T>);   // bytes per value: 16 (FP64) or 8 (FP32)
    constexpr int VPL = 16 / VB;             // values one lane loads (16 bytes)
    constexpr int LPR = LPB / VPL;           // lanes per slot row
    constexpr int RPI = 64 / LPR;            // rows of a slot one wave instruction fills
    constexpr int IPW = M / (RPI * NW);      // slot loads per wave (and per lane) per field
    constexpr unsigned ROWB = LPB * VB;      // bytes per slot row
    static_assert(R0 == E && (E == 4 || E == 8) && P::nb(0) == 1 && P::full(0) && P::radix(P::S - 1) == R0 &&
                      P::nb(P::S - 1) == 1,
                  "the slot reads assume one butterfly of all a thread's values in the first and last stage");
    static_assert((E == 8) == (RT == 4 || RT == 5),
                  "eight values per thread: the one-exchange transforms zw8_inv, zw8_fwd, zs_inv, zs_fwd");
    static_assert(NTH % 64 == 0 && LPB % VPL == 0 && 64 % LPR == 0 && M % (RPI * NW) == 0 &&
                      (R0 - 1) * Q * ROWB < 65536,
                  "slot geometry");
    __shared__ C2<T> slot[3][M * LPB];
    // the FFT exchange: split re / im arrays with padded rows, or (IL)
    // interleaved complex values, rows of LPB columns
    __shared__ __align__(16) T sre[IL ? 2 * M * LPB : M * CS];
    __shared__ T sim[IL ? 1 : M * CS];
    C2<T>* const sx = reinterpret_cast<C2<T>*>(sre);
    __shared__ C2<T> tw[M];
    __shared__ C2<T> twn[HALF ? M : 1];
    W += blockIdx.y * boff;
    for (int t = threadIdx.x; t < M; t += NTH) {
        tw[t] = twg[t];
        if constexpr (HALF) twn[t] = twng[t];
    }
    __syncthreads();
    const int c = threadIdx.x % LPB;
    const int tid = threadIdx.x / LPB;
    const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
    const int tiles_x = (ncx + LPB - 1) / LPB;
    const unsigned es = (unsigned)nx * (unsigned)ny;
    const unsigned uF = (unsigned)F;
    static_assert(RT == 0 || IL, "register twiddles: interleaved exchange");
    static_assert(RT != 3 || (M == 128 && P::TPL == 32 && E == 4 && LPB == 8),
                  "the one-exchange transforms: 128-point lines, eight per block of four waves");
    static_assert(RT != 4 || (M == 128 && P::TPL == 16 && LPB == 16 && HALF),
                  "the one-exchange transforms of eight values: 128-point lines, sixteen per block of four waves");
    static_assert(RT != 5 || (M == 256 && P::TPL == 32 && LPB == 16 && HALF),
                  "the one-exchange transforms of eight values: 256-point lines, sixteen per block of eight waves");
    // RT >= 1: this thread's packing factors e^{-2 pi i k/nz}, k = tid + r Q
    // (the positions of its first-stage inputs and of its last-stage outputs);
    // RT == 2: also its stage twiddles, loaded once for every tile; RT == 3:
    // the transforms with one LDS exchange (zw_inv, zw_fwd) and their twiddles;
    // RT == 4: the same with eight values per thread (zw8_inv, zw8_fwd); RT ==
    // 5: for 256-point lines (zs_inv, zs_fwd)
    C2<T> twp[RT >= 1 ? E : 1], tws[RT == 2 ? rt_size<P>() : 1];
    ZwTw<T> zt;
    ZwTw8<T> zt8;
    ZsTw<T> zts;
    if constexpr (RT >= 1 && HALF) {
#pragma unroll
        for (int r = 0; r < E; r++) twp[r] = twng[tid + r * Q];
    }
    if constexpr (RT == 2) rt_load<P, 0, T>(tws, tid, twg);
    if constexpr (RT == 3) zt.load(twg, tid);
    if constexpr (RT == 4) zt8.load(twg, tid);
    if constexpr (RT == 5) zts.load(twg, tid);
    auto idx = [&](int pos) { return IL ? pos * LPB + c : pos * CS + c; };
    auto fft = [&](auto sign, C2<T>* v) {
        constexpr int SG = decltype(sign)::value;
        if constexpr (RT == 5) {
            if constexpr (SG > 0)
                zs_inv<T>(v, zts, sx, wave, lane);
            else
                zs_fwd<T>(v, zts, sx, wave, lane);
        } else if constexpr (RT == 4) {
            if constexpr (SG > 0)
                zw8_inv<T>(v, zt8, sx, wave, lane);
            else
                zw8_fwd<T>(v, zt8, sx, wave, lane);
        } else if constexpr (RT == 3) {
            if constexpr (SG > 0)
                zw_inv<T>(v, zt, sx, wave, lane);
            else
                zw_fwd<T>(v, zt, sx, wave, lane);
        } else if constexpr (RT == 2)
            line_fft_cr<P, SG, T, decltype(idx), SyncLds>(v, tid, sx, idx, tws);
        else if constexpr (IL)
            line_fft_c<P, SG, T, decltype(idx), SyncLds>(v, tid, sx, idx, tw);
        else
            line_fft<P, SG, T, decltype(idx), SyncLds>(v, tid, sre, sim, idx, tw);
    };
    // field f of tile t into slot sl: lane -> (columns VPL (lane % LPR) and on,
    // row lane / LPR) of each RPI-row group; the LDS destination of a wave
    // instruction is uniform (M0) and the lanes land 16 bytes apart
    auto fetch = [&](int t, int f, int sl) {
        const int j = t / tiles_x;
        const int i = (t % tiles_x) * LPB + (lane % LPR) * VPL;
        const unsigned cb = i < ncx ? (unsigned)i + (unsigned)j * (unsigned)nx : 0u;
#pragma unroll
        for (int k = 0; k < IPW; k++) {
            const int row0 = (wave * IPW + k) * RPI;
            const C2<T>* src = W + (cb + (unsigned)f * uF + (unsigned)(row0 + lane / LPR) * es);
            __builtin_amdgcn_global_load_lds((const __attribute__((address_space(1))) void*)src,
                                             (__attribute__((address_space(3))) void*)&slot[sl][row0 * LPB], 16,
                                             0, LAUX);
        }
    };
    // this thread's four values of a slot, rows tid + r Q from LDS address a,
    // and (read8) their mirrors, rows (Q - tid) + (3 - r) Q from address m, in
    // one asm block that waits for them
    typedef T v2 __attribute__((ext_vector_type(2)));
#define SH_ZD_READ(INS)                                                                              \
    asm volatile(INS " %0, %4\n\t" INS " %1, %4 offset:%5\n\t" INS " %2, %4 offset:%6\n\t" INS      \
                     " %3, %4 offset:%7\n\t"                                                           \
                     "s_waitcnt lgkmcnt(0)"                                                         \
                 : "=&v"(o[0]), "=&v"(o[1]), "=&v"(o[2]), "=&v"(o[3])                              \
                 : "v"(a), "n"(Q * ROWB), "n"(2 * Q * ROWB), "n"(3 * Q * ROWB)                      \
                 : "memory")
#define SH_ZD_READ8(INS)                                                                             \
    asm volatile(INS " %0, %8\n\t" INS " %1, %8 offset:%10\n\t" INS " %2, %8 offset:%11\n\t" INS    \
                     " %3, %8 offset:%12\n\t" INS " %4, %9 offset:%12\n\t" INS " %5, %9 offset:%11\n\t" \
                 INS " %6, %9 offset:%10\n\t" INS " %7, %9\n\t"                                      \
                     "s_waitcnt lgkmcnt(0)"                                                         \
                 : "=&v"(o[0]), "=&v"(o[1]), "=&v"(o[2]), "=&v"(o[3]), "=&v"(mo[0]), "=&v"(mo[1]),   \
                   "=&v"(mo[2]), "=&v"(mo[3])                                                       \
                 : "v"(a), "v"(m), "n"(Q * ROWB), "n"(2 * Q * ROWB), "n"(3 * Q * ROWB)              \
                 : "memory")
    auto read4 = [](unsigned a, v2* o) {
        if constexpr (VB == 16)
            SH_ZD_READ("ds_read_b128");
        else
            SH_ZD_READ("ds_read_b64");
    };
    auto read8 = [](unsigned a, unsigned m, v2* o, v2* mo) {
        if constexpr (VB == 16)
            SH_ZD_READ8("ds_read_b128");
        else
            SH_ZD_READ8("ds_read_b64");
    };
#undef SH_ZD_READ
#undef SH_ZD_READ8
    // X[k] and X[M-k] of this thread's first-stage positions k = tid + r Q
    // from slot sl (X[M] = 0, the Nyquist plane), packed into Z and
    // inverse-transformed, scaled by 1/(Nx Ny Nz)
    auto inv = [&](int sl, C2<T>* x) {
        const unsigned base = lds_offset(&slot[sl][0]);
        const unsigned own = base + (unsigned)(tid * LPB + c) * VB;
        v2 av[E];
        if constexpr (!HALF) {
            // the full spectrum: this thread's first-stage values as they are
            read4(own, av);
#pragma unroll
            for (int r = 0; r < 4; r++) x[r] = mk<T>(av[r].x, av[r].y);
        } else {
            const unsigned mir = base + (unsigned)((tid == 0 ? Q : Q - tid) * LPB + c) * VB;
            v2 mv[E];  // X[M - k] for r = 0..E-1
            if constexpr (E == 8) {
                // rows tid + r Q (r = 0..7) and their mirrors (Q - tid) + (7 - r) Q
                read8(own, mir + 4u * Q * ROWB, av, mv);
                read8(own + 4u * Q * ROWB, mir, av + 4, mv + 4);
            } else {
                read8(own, mir, av, mv);
            }
#pragma unroll
            for (int r = 0; r < E; r++) {
                const int pos = tid + r * Q;
                const C2<T> a = mk<T>(av[r].x, av[r].y);
                const C2<T> b = pos == 0 ? mk<T>(T(0), T(0)) : mk<T>(mv[r].x, mv[r].y);
                const C2<T> sum = mk<T>(a.x + b.x, a.y - b.y);  // X[k] + conj X[M-k]
                const C2<T> dif = mk<T>(a.x - b.x, a.y + b.y);  // X[k] - conj X[M-k]
                const C2<T> w = RT >= 1 ? twp[r] : twn[pos];    // e^{-2 pi i k/nz}
                const T tx = w.x * dif.x + w.y * dif.y;         // t = conj(w) dif
                const T ty = w.x * dif.y - w.y * dif.x;
                x[r] = mk<T>(sum.x - ty, sum.y + tx);           // Z = sum + i t
            }
        }
        SyncLds::run();
        fft(std::integral_constant<int, 1>{}, x);
#pragma unroll
        for (int s2 = 0; s2 < E; s2++) x[s2] = rmul<T>(scale, x[s2]);
    };
    // the product of two fields: real and imaginary parts separately (two grid
    // points of the packed half spectrum) or the full complex product
    auto pmul = [](C2<T> a, C2<T> b) {
        if constexpr (HALF)
            return mk<T>(a.x * b.x, a.y * b.y);
        else
            return cmul<T>(a, b);
    };
    // fields in the order w0 w1 u2 w2 u0 u1 (3 4 2 5 0 1), slots 0 1 2 0 1 2
    int t = blockIdx.x;
    if (t < ntile) {
        fetch(t, 3, 0);
        fetch(t, 4, 1);
        fetch(t, 2, 2);
    }
    for (; t < ntile; t += gridDim.x) {
        const int tn = t + gridDim.x;
        const bool more = tn < ntile;
        const int j = t / tiles_x;
        const int i = (t % tiles_x) * LPB + c;
        const bool valid = i < ncx;
        const unsigned cb = valid ? (unsigned)i + (unsigned)j * (unsigned)nx : 0u;
        C2<T> w[3][E], o[3][E], x[E];
        // at each wait, the loads issued after the awaited field's are those of
        // the next two fields (and, for this tile's first three, the previous
        // tile's stores, which the wait may drain as well)
        vm_wait_barrier<2 * IPW>();
        inv(0, w[0]);
        fetch(t, 5, 0);
        vm_wait_barrier<2 * IPW>();
        inv(1, w[1]);
        fetch(t, 0, 1);
        vm_wait_barrier<2 * IPW>();
        inv(2, x);  // u2
#pragma unroll
        for (int s2 = 0; s2 < E; s2++) {
            o[0][s2] = csub<T>(mk<T>(T(0), T(0)), pmul(x[s2], w[1][s2]));
            o[1][s2] = pmul(x[s2], w[0][s2]);
        }
        fetch(t, 1, 2);
        vm_wait_barrier<2 * IPW>();
        inv(0, w[2]);
        if (more) fetch(tn, 3, 0);
        if (more)
            vm_wait_barrier<2 * IPW>();
        else
            vm_wait_barrier<IPW>();
        inv(1, x);  // u0
#pragma unroll
        for (int s2 = 0; s2 < E; s2++) {
            o[1][s2] = csub<T>(o[1][s2], pmul(x[s2], w[2][s2]));
            o[2][s2] = pmul(x[s2], w[1][s2]);
        }
        if (more) fetch(tn, 4, 1);
        if (more)
            vm_wait_barrier<2 * IPW>();
        else
            vm_wait_barrier<0>();
        inv(2, x);  // u1
#pragma unroll
        for (int s2 = 0; s2 < E; s2++) {
            o[0][s2] = cadd<T>(o[0][s2], pmul(x[s2], w[2][s2]));
            o[2][s2] = csub<T>(o[2][s2], pmul(x[s2], w[0][s2]));
        }
        if (more) fetch(tn, 2, 2);
#pragma unroll
        for (int d = 0; d < 3; d++) {
#pragma unroll
            for (int s2 = 0; s2 < E; s2++) x[s2] = o[d][s2];
            SyncLds::run();
            fft(std::integral_constant<int, -1>{}, x);
            if constexpr (!HALF) {
                if (valid)
                    for_out<P>(tid, [&](int s2, int pos) {
                        st_line<NT, T>(W + (cb + (unsigned)d * uF + (unsigned)pos * es), x[s2]);
                    });
                continue;
            }
            // Z -> X (Z[M-k] from another thread through LDS)
            SyncLds::run();
            for_out<P>(tid, [&](int s2, int pos) {
                if constexpr (IL) {
                    sx[idx(pos)] = x[s2];
                } else {
                    sre[idx(pos)] = x[s2].x;
                    sim[idx(pos)] = x[s2].y;
                }
            });
            SyncLds::run();
            if (valid)
                for_out<P>(tid, [&](int s2, int pos) {
                    const C2<T> a = x[s2];
                    const int mp = pos == 0 ? 0 : M - pos;
                    const C2<T> b = IL ? sx[idx(mp)] : mk<T>(sre[idx(mp)], sim[idx(mp)]);
                    const C2<T> sum = mk<T>(a.x + b.x, a.y - b.y);  // Z[k] + conj Z[M-k]
                    const C2<T> dif = mk<T>(a.x - b.x, a.y + b.y);  // Z[k] - conj Z[M-k]
                    const C2<T> wv = RT >= 1 ? twp[s2] : twn[pos];  // last stage: s2 = r, pos = tid + r Q
                    const T tx = wv.x * dif.x - wv.y * dif.y;  // t = w dif
                    const T ty = wv.x * dif.y + wv.y * dif.x;
                    st_line<NT, T>(W + (cb + (unsigned)d * uF + (unsigned)pos * es),
                                   mk<T>(T(0.5) * (sum.x + ty), T(0.5) * (sum.y - tx)));
                });
        }
    }
}

// expand a half-spectrum field to the full one in place: planes kz > nz/2 from
// f(-k) = conj f(k), the Nyquist plane kz = nz/2 zeroed (count fields at fstride)
template <typename T>
__global__ void k_expand_half(C2<T>* __restrict__ f, int count, long fstride, int nx, int ny, int nz) {
    const int h = nz / 2;
    const long plane = (long)nx * ny;
    const long nup = plane * (nz - h);  // planes h .. nz-1
    for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < nup * count; e += (long)gridDim.x * blockDim.x) {
        const int fld = (int)(e / nup);
        const long r = e % nup;
        const int k = h + (int)(r / plane);
        const int j = (int)((r % plane) / nx), i = (int)(r % nx);
        C2<T>* b = f + fld * fstride;
        C2<T> v = mk<T>(T(0), T(0));
        if (k != h) {
            const C2<T> m = b[(long)((nx - i) % nx) + (long)nx * ((ny - j) % ny) + plane * (nz - k)];
            v = mk<T>(m.x, -m.y);
        }
        b[(long)i + (long)nx * j + plane * k] = v;
    }
}

// per-block maxima of |f(k) - conj f(-k)| and |f(k)| over count fields (FP64)
__global__ __launch_bounds__(256) void k_herm_dev(const double2* __restrict__ f, int count, long fstride, int nx,
                                                  int ny, int nz, double* __restrict__ part) {
    __shared__ double sd[256], sm[256];
    const long plane = (long)nx * ny, N = plane * nz;
    double dmax = 0, fmax = 0;
    for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < N * count; e += (long)gridDim.x * blockDim.x) {
        const int fld = (int)(e / N);
        const long r = e % N;
        const int k = (int)(r / plane), j = (int)((r % plane) / nx), i = (int)(r % nx);
        const double2* b = f + fld * fstride;
        const double2 a = b[r];
        const double2 m = b[(long)((nx - i) % nx) + (long)nx * ((ny - j) % ny) + plane * ((nz - k) % nz)];
        dmax = fmax_d(dmax, hypot(a.x - m.x, a.y + m.y));
        fmax = fmax_d(fmax, hypot(a.x, a.y));
    }
    sd[threadIdx.x] = dmax;
    sm[threadIdx.x] = fmax;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            sd[threadIdx.x] = fmax_d(sd[threadIdx.x], sd[threadIdx.x + s]);
            sm[threadIdx.x] = fmax_d(sm[threadIdx.x], sm[threadIdx.x + s]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = sd[0];
        part[2 * blockIdx.x + 1] = sm[0];
    }
}

// E: x-pass forward of the products (-> N1, Nyquist zeroed) and the
// Fourier-space update of exec/ns.cpp:171-187; writes V1 <- V, N0 <- N1 and
// per-block sums of |div V| (exec/ns.cpp:209-210).  With EMIT it also runs
// the next step's first pass: the inverse x transform of the new V and of
// omega = i k x V (exec/ns.cpp:206, 216) into W[0..6), which kernel A would
// otherwise do (same slot->position map, so no reshuffle is needed).  The
// products are read from W[0..3) before any line of W is rewritten, and a
// block only rewrites its own lines.
// BL: the line's V and N0 stores are issued after the whole update, so no
// slot's V1/N0 loads wait behind an earlier slot's stores (vector-memory
// counters count loads and stores together)
// YS: the merged AB2 history (see k_ns_fourier_xd): N0 holds Y = (V1 + lap1)
// - K/2 N0, V is neither read nor written, and the new V survives only as the
// emitted x-inverse (so YS needs EMIT).  E5 (with YS): five emitted fields,
// the x-inverses of V and of i kx Vy, i kx Vz (see k_ns_y_curl).  LNT:
// nontemporal loads (every input is read once)
template <typename T, class P, int LPB, bool EMIT, bool BL = false, bool NT = false, bool YS = false,
          bool E5 = false, bool LNT = false>
__global__ __launch_bounds__(P::TPL* LPB) void k_ns_fourier_x(C2<T>* __restrict__ W,
                                                              C2<T>* __restrict__ V,
                                                              C2<T>* __restrict__ N0,
                                                              const C2<T>* __restrict__ frc,
                                                              const C2<T>* __restrict__ tw, int ny,
                                                              int nz, int nzp, long F,
                                                              const StepCoef<T>* __restrict__ cfs,
                                                              double* __restrict__ divpart,
                                                              long boff, int divstride) {
    static_assert(!EMIT || (P::radix(0) == P::radix(P::S - 1) && P::nb(0) == P::nb(P::S - 1)),
                  "emitting the next inverse pass needs a symmetric plan");
    static_assert(!YS || EMIT, "the merged history keeps V only as the emitted x-inverse");
    static_assert(!E5 || YS, "five emitted fields with the merged history only");
    // per-box coefficients (each box of an ensemble has its own viscosity)
    const StepCoef<T> cf = cfs[blockIdx.y];
    W += blockIdx.y * boff;
    V += blockIdx.y * boff;
    N0 += blockIdx.y * boff;
    const bool forced = frc != nullptr;  // body force added to the nonlinear term
    if (forced) frc += blockIdx.y * boff;
    divpart += (long)blockIdx.y * divstride;
    constexpr int N = P::N;
    constexpr int LS = LdsLine<N>::LS;
    constexpr int E = P::E;
    __shared__ T sre[LPB * LS];
    __shared__ T sim[LPB * LS];
    const int tid = threadIdx.x % P::TPL;
    const int l = threadIdx.x / P::TPL;
    const long nlines = (long)ny * nzp;
    const long line = (long)blockIdx.x * LPB + l;
    const bool valid = line < nlines;
    const long lc = valid ? line : 0;
    const int jj = (int)(lc % ny), kk = (int)(lc / ny);
    const T k1 = freq<T>(jj, ny), k2 = freq<T>(kk, nz);
    const bool zline = (jj == ny / 2) || (kk == nz / 2);
    // half-spectrum layout: a stored plane kz > 0 also stands for the plane -kz,
    // whose |div V| values are the same (f(-k) = conj f(k))
    const double dw = (nzp < nz && kk > 0) ? 2.0 : 1.0;
    const long off = lc * N;
    auto idx = [&](int pos) { return l * LS + pos + (pos >> 4); };
    auto ld = [](const C2<T>* a) {
        if constexpr (LNT) {
            typedef T t2 __attribute__((ext_vector_type(2)));
            const t2 w = __builtin_nontemporal_load(reinterpret_cast<const t2*>(a));
            return mk<T>(w.x, w.y);
        } else {
            return *a;
        }
    };
    auto load = [&](int f, C2<T>* x) { for_in<P>(tid, [&](int s, int pos) { x[s] = ld(W + f * F + off + pos); }); };
    // forward x transforms of the three products; the next field's loads are
    // issued before the current transform so they overlap its LDS phases
    C2<T> n1[3][E], x[E], pf[E];
    load(0, x);
    load(1, pf);
    line_fft<P, -1, T>(x, tid, sre, sim, idx, tw);
#pragma unroll
    for (int s = 0; s < E; s++) n1[0][s] = x[s];
#pragma unroll
    for (int s = 0; s < E; s++) x[s] = pf[s];
    load(2, pf);
    __syncthreads();
    line_fft<P, -1, T>(x, tid, sre, sim, idx, tw);
#pragma unroll
    for (int s = 0; s < E; s++) n1[1][s] = x[s];
    __syncthreads();
    line_fft<P, -1, T>(pf, tid, sre, sim, idx, tw);
#pragma unroll
    for (int s = 0; s < E; s++) n1[2][s] = pf[s];
    double dsum = 0;
    C2<T> vn[3][E];
#pragma unroll
    for (int d = 0; d < 3; d++)
#pragma unroll
        for (int s = 0; s < E; s++) vn[d][s] = mk<T>(T(0), T(0));
    C2<T> nna[BL ? 3 : 1][BL ? E : 1];
    if (valid) {
        for_out<P>(tid, [&](int s, int pos) {
            const T k0 = freq<T>(pos, N);
            const bool z = zline || pos == N / 2 ||
                           (cf.dealias == 1 && (3 * fabs(k0) > N || 3 * fabs(k1) > ny || 3 * fabs(k2) > nz));
            const T sig = cf.dealias == 2 ? expfilter<T>(k0, N) * expfilter<T>(k1, ny) * expfilter<T>(k2, nz) : T(1);
            C2<T> nn[3], v1[3], f[3], L[3], S[3];
#pragma unroll
            for (int d = 0; d < 3; d++) {
                nn[d] = z ? mk<T>(T(0), T(0)) : forced ? cadd<T>(n1[d][s], frc[d * F + off + pos]) : n1[d][s];
                if (cf.dealias == 2) nn[d] = rmul<T>(sig, nn[d]);
                if constexpr (YS) {
                    // S = Y + K/2*(3.*cross1)  (= (V1 + lap1) + F)
                    S[d] = cadd<T>(ld(N0 + d * F + off + pos), rmul<T>(cf.Kh, rmul<T>(T(3), nn[d])));
                } else {
                    v1[d] = ld(V + d * F + off + pos);
                    const C2<T> n0 = ld(N0 + d * F + off + pos);
                    // F = K/2*(3.*cross1 - cross0)
                    f[d] = rmul<T>(cf.Kh, csub<T>(rmul<T>(T(3), nn[d]), n0));
                }
            }
            // lap1 = nu*K/2*V1.laplacian() = c1 * (-(k.k) V1)
            const T q = -(k0 * k0 + k1 * k1 + k2 * k2);
            if constexpr (!YS) {
#pragma unroll
                for (int d = 0; d < 3; d++) {
                    L[d] = rmul<T>(cf.c1, rmul<T>(q, v1[d]));
                    S[d] = cadd<T>(cadd<T>(v1[d], L[d]), f[d]);  // (V1 + lap1) + F
                }
            }
            // G = div(S); G.laplacian_inverse() (a=0, b=1)
            C2<T> G = mk<T>(T(0), T(0));
            G = cadd<T>(G, imul<T>(k0, S[0]));
            G = cadd<T>(G, imul<T>(k1, S[1]));
            G = cadd<T>(G, imul<T>(k2, S[2]));
            T sum = 0;
            sum += k0 * k0;
            sum += k1 * k1;
            sum += k2 * k2;
            // one reciprocal per point for each factor, shared by the components
            // (within an ulp of dividing each component)
            const T rg = sum != 0 ? T(1) / (T(0) - T(1) * sum) : T(1);
            G.x *= rg;
            G.y *= rg;
            const double rv = 1.0 / (1.0 - (-cf.c1d) * (double)sum);  // 1 where sum == 0
            const T kk3[3] = {k0, k1, k2};
#pragma unroll
            for (int d = 0; d < 3; d++) {
                C2<T> vv;
                if constexpr (YS) {
                    vv = csub<T>(S[d], imul<T>(kk3[d], G));  // Vhalf + lap1 = S - G.gradient()
                } else {
                    // Vhalf = V1 + F - G.gradient();  V = Vhalf + lap1 (exec/ns2.cpp:185: V = Vhalf)
                    C2<T> vh = csub<T>(cadd<T>(v1[d], f[d]), imul<T>(kk3[d], G));
                    vv = cf.be ? vh : cadd<T>(vh, L[d]);
                }
                // V[d].laplacian_inverse(1, -nu*K/2): divide by 1 + (nu K/2)|k|^2 (FP64 factor)
                vv = mk<T>(T((double)vv.x * rv), T((double)vv.y * rv));
                vn[d][s] = vv;
            }
            C2<T> dv = mk<T>(T(0), T(0));
            dv = cadd<T>(dv, imul<T>(k0, vn[0][s]));
            dv = cadd<T>(dv, imul<T>(k1, vn[1][s]));
            dv = cadd<T>(dv, imul<T>(k2, vn[2][s]));
            dsum += dw * sqrt((double)dv.x * (double)dv.x + (double)dv.y * (double)dv.y);  // |div V| (no overflow here)
#pragma unroll
            for (int d = 0; d < 3; d++) {
                // YS: the next step's Y = (V + lap V) - K/2 cross1 in place of N0
                const C2<T> keep = YS ? csub<T>(cadd<T>(vn[d][s], rmul<T>(cf.c1, rmul<T>(q, vn[d][s]))),
                                                rmul<T>(cf.Kh, nn[d]))
                                      : nn[d];
                if constexpr (BL) {
                    nna[d][s] = keep;
                } else {
                    if constexpr (!YS) st_line<NT, T>(V + d * F + off + pos, vn[d][s]);
                    st_line<NT, T>(N0 + d * F + off + pos, keep);
                }
            }
        });
        if constexpr (BL)
            for_out<P>(tid, [&](int s, int pos) {
#pragma unroll
                for (int d = 0; d < 3; d++) {
                    if constexpr (!YS) st_line<NT, T>(V + d * F + off + pos, vn[d][s]);
                    st_line<NT, T>(N0 + d * F + off + pos, nna[d][s]);
                }
            });
    }
    if constexpr (EMIT) {
        // next step's kernel A: inverse x of V (g = 0..2) and of i k x V (g = 3..5);
        // E5: of i kx Vy and i kx Vz (g = 3, 4)
#pragma unroll
        for (int g = 0; g < (E5 ? 5 : 6); g++) {
            for_out<P>(tid, [&](int s, int pos) {
                const T k0 = freq<T>(pos, N);
                if (g < 3)
                    x[s] = vn[g][s];
                else if (E5)
                    x[s] = imul<T>(k0, vn[g - 2][s]);
                else if (g == 3)
                    x[s] = csub<T>(imul<T>(k1, vn[2][s]), imul<T>(k2, vn[1][s]));
                else if (g == 4)
                    x[s] = csub<T>(imul<T>(k2, vn[0][s]), imul<T>(k0, vn[2][s]));
                else
                    x[s] = csub<T>(imul<T>(k0, vn[1][s]), imul<T>(k1, vn[0][s]));
            });
            __syncthreads();
            line_fft<P, 1, T>(x, tid, sre, sim, idx, tw);
            if (valid) for_out<P>(tid, [&](int s, int pos) { st_line<NT, T>(W + g * F + off + pos, x[s]); });
        }
    }
    // block sum of |div V| (FP64), deterministic; reuses the FFT's LDS
    static_assert(LPB * LS * sizeof(T) >= P::TPL * LPB * sizeof(double), "LDS too small for reduction");
    __syncthreads();
    double* red = reinterpret_cast<double*>(sre);
    red[threadIdx.x] = dsum;
    __syncthreads();
    for (int n = blockDim.x; n > 1;) {
        const int h = (n + 1) / 2;
        if ((int)threadIdx.x < n - h) red[threadIdx.x] += red[threadIdx.x + h];
        __syncthreads();
        n = h;
    }
    if (threadIdx.x == 0) divpart[blockIdx.x] = red[0];
}

// A one-wave line transform whose LDS exchanges are inline asm: interleaved
// complex values (one ds_write/ds_read_b128 each) at the swizzled index
// SwzIdx of a buffer of N values at LDS address buf, and the twiddles from a
// stage-major table at LDS address twb.  The compiler sees no LDS access in
// it, so it inserts no wait for loads into LDS in flight meanwhile (which it
// would drain at once, vmcnt(0), before any access it cannot tell apart from
// their targets); the caller orders the buffer against those loads.  The
// block is one wave, whose LDS instructions execute in issue order, so a put
// needs no wait before the get that follows it; each get waits for its reads.
typedef double xa_d2 __attribute__((ext_vector_type(2)));

template <class P, int s>
__device__ __forceinline__ void xa_put(const C2<double>* x, int tid, unsigned buf) {
    constexpr int R = P::radix(s), Ns = P::ns(s);
    const int base = (tid / Ns) * Ns * R + (tid % Ns);
#pragma unroll
    for (int r = 0; r < R; r++) {
        const unsigned a = buf + 16u * (unsigned)SwzIdx()(base + r * Ns);
        const xa_d2 v = {x[r].x, x[r].y};
        asm volatile("ds_write_b128 %0, %1" ::"v"(a), "v"(v) : "memory");
    }
}

// the values tid + r N/4 (r = 0..3) of radix-4 stage s and, in the same asm
// block (one wait for both), this lane's twiddles of that stage (k = tid mod Ns)
template <class P, int s>
__device__ __forceinline__ void xa_get(C2<double>* x, C2<double>* w, int tid, unsigned buf, unsigned twb) {
    constexpr int Q = P::N / 4;
    constexpr int D = 16 * Q;
    constexpr int Ns = P::ns(s);
    static_assert(Q % 64 == 0 && 3 * D < 65536 && 32 * Ns < 65536,
                  "constant offsets between a lane's four values and three twiddles");
    const unsigned a = buf + 16u * (unsigned)SwzIdx()(tid);
    const unsigned t = twb + 16u * (unsigned)(sm_off<P, s>() + tid % Ns);
    xa_d2 v0, v1, v2, v3, w1, w2, w3;
    asm volatile(
        "ds_read_b128 %0, %7\n\t"
        "ds_read_b128 %1, %7 offset:%9\n\t"
        "ds_read_b128 %2, %7 offset:%10\n\t"
        "ds_read_b128 %3, %7 offset:%11\n\t"
        "ds_read_b128 %4, %8\n\t"
        "ds_read_b128 %5, %8 offset:%12\n\t"
        "ds_read_b128 %6, %8 offset:%13\n\t"
        "s_waitcnt lgkmcnt(0)"
        : "=&v"(v0), "=&v"(v1), "=&v"(v2), "=&v"(v3), "=&v"(w1), "=&v"(w2), "=&v"(w3)
        : "v"(a), "v"(t), "n"(D), "n"(2 * D), "n"(3 * D), "n"(16 * Ns), "n"(32 * Ns)
        : "memory");
    x[0] = mk<double>(v0.x, v0.y);
    x[1] = mk<double>(v1.x, v1.y);
    x[2] = mk<double>(v2.x, v2.y);
    x[3] = mk<double>(v3.x, v3.y);
    w[0] = mk<double>(w1.x, w1.y);
    w[1] = mk<double>(w2.x, w2.y);
    w[2] = mk<double>(w3.x, w3.y);
}

// radix-4 butterfly of a stage after the first: twiddles w (conjugated for the
// inverse sign), then the 4-point DFT -- stage_bfly's arithmetic
template <int SIGN>
__device__ __forceinline__ void xa_bfly(C2<double>* x, const C2<double>* w) {
#pragma unroll
    for (int r = 1; r < 4; r++) {
        C2<double> t = w[r - 1];
        if constexpr (SIGN > 0) t.y = -t.y;
        x[r] = cmul<double>(x[r], t);
    }
    Dft<double, 4, SIGN>::run(x);
}

// tws: the stage-major twiddle table (sm_fill)
template <class P, int SIGN>
__device__ __forceinline__ void line_fft_xa(C2<double>* x, int tid, unsigned buf,
                                            const C2<double>* __restrict__ tws) {
    static_assert(P::TPL == 64 && P::S == 4 && P::radix(0) == 4 && P::radix(1) == 4 && P::radix(2) == 4 &&
                      P::radix(3) == 4 && P::N == 256 && P::ns(0) == 1,
                  "four radix-4 stages, one butterfly per lane");
    const unsigned twb = lds_offset(tws);
    C2<double> w[3];
    stage_bfly<P, 0, SIGN, double>(x, tid, tws);  // no twiddles in the first stage
    xa_put<P, 0>(x, tid, buf);
    xa_get<P, 1>(x, w, tid, buf, twb);
    xa_bfly<SIGN>(x, w);
    xa_put<P, 1>(x, tid, buf);
    xa_get<P, 2>(x, w, tid, buf, twb);
    xa_bfly<SIGN>(x, w);
    xa_put<P, 2>(x, tid, buf);
    xa_get<P, 3>(x, w, tid, buf, twb);
    xa_bfly<SIGN>(x, w);
}

// Three lines transformed together by one wave, stage by stage (three times
// the independent arithmetic between LDS round trips): line d exchanges in
// the buffer at LDS address buf + d * 16 N, all with the same twiddles.
template <class P, int s>
__device__ __forceinline__ void xa_put3(const C2<double> (&x)[3][4], int tid, unsigned buf) {
    constexpr int R = P::radix(s), Ns = P::ns(s);
    const int base = (tid / Ns) * Ns * R + (tid % Ns);
#pragma unroll
    for (int r = 0; r < R; r++) {
        const unsigned a = buf + 16u * (unsigned)SwzIdx()(base + r * Ns);
#pragma unroll
        for (int d = 0; d < 3; d++) {
            const xa_d2 v = {x[d][r].x, x[d][r].y};
            asm volatile("ds_write_b128 %0, %1 offset:%2" ::"v"(a), "v"(v), "n"(d * 16 * P::N) : "memory");
        }
    }
}

template <class P, int s>
__device__ __forceinline__ void xa_get3(C2<double> (&x)[3][4], C2<double>* w, int tid, unsigned buf,
                                        unsigned twb) {
    constexpr int Q = P::N / 4;
    constexpr int D = 16 * Q;
    constexpr int Ns = P::ns(s);
    static_assert(Q % 64 == 0 && 2 * 16 * P::N + 3 * D < 65536 && 32 * Ns < 65536, "constant offsets");
    const unsigned a0 = buf + 16u * (unsigned)SwzIdx()(tid);
    const unsigned a1 = a0 + 16u * P::N, a2 = a1 + 16u * P::N;
    const unsigned t = twb + 16u * (unsigned)(sm_off<P, s>() + tid % Ns);
    xa_d2 v[3][4], w1, w2, w3;
    asm volatile(
        "ds_read_b128 %0, %15\n\t"
        "ds_read_b128 %1, %15 offset:%19\n\t"
        "ds_read_b128 %2, %15 offset:%20\n\t"
        "ds_read_b128 %3, %15 offset:%21\n\t"
        "ds_read_b128 %4, %16\n\t"
        "ds_read_b128 %5, %16 offset:%19\n\t"
        "ds_read_b128 %6, %16 offset:%20\n\t"
        "ds_read_b128 %7, %16 offset:%21\n\t"
        "ds_read_b128 %8, %17\n\t"
        "ds_read_b128 %9, %17 offset:%19\n\t"
        "ds_read_b128 %10, %17 offset:%20\n\t"
        "ds_read_b128 %11, %17 offset:%21\n\t"
        "ds_read_b128 %12, %18\n\t"
        "ds_read_b128 %13, %18 offset:%22\n\t"
        "ds_read_b128 %14, %18 offset:%23\n\t"
        "s_waitcnt lgkmcnt(0)"
        : "=&v"(v[0][0]), "=&v"(v[0][1]), "=&v"(v[0][2]), "=&v"(v[0][3]), "=&v"(v[1][0]), "=&v"(v[1][1]),
          "=&v"(v[1][2]), "=&v"(v[1][3]), "=&v"(v[2][0]), "=&v"(v[2][1]), "=&v"(v[2][2]), "=&v"(v[2][3]),
          "=&v"(w1), "=&v"(w2), "=&v"(w3)
        : "v"(a0), "v"(a1), "v"(a2), "v"(t), "n"(D), "n"(2 * D), "n"(3 * D), "n"(16 * Ns), "n"(32 * Ns)
        : "memory");
#pragma unroll
    for (int d = 0; d < 3; d++)
#pragma unroll
        for (int r = 0; r < 4; r++) x[d][r] = mk<double>(v[d][r].x, v[d][r].y);
    w[0] = mk<double>(w1.x, w1.y);
    w[1] = mk<double>(w2.x, w2.y);
    w[2] = mk<double>(w3.x, w3.y);
}

template <class P, int SIGN>
__device__ __forceinline__ void line_fft_xa3(C2<double> (&x)[3][4], int tid, unsigned buf,
                                             const C2<double>* __restrict__ tws) {
    static_assert(P::TPL == 64 && P::S == 4 && P::radix(0) == 4 && P::radix(1) == 4 && P::radix(2) == 4 &&
                      P::radix(3) == 4 && P::N == 256 && P::ns(0) == 1,
                  "four radix-4 stages, one butterfly per lane");
    const unsigned twb = lds_offset(tws);
    C2<double> w[3];
#pragma unroll
    for (int d = 0; d < 3; d++) stage_bfly<P, 0, SIGN, double>(x[d], tid, tws);  // no twiddles
    xa_put3<P, 0>(x, tid, buf);
    xa_get3<P, 1>(x, w, tid, buf, twb);
#pragma unroll
    for (int d = 0; d < 3; d++) xa_bfly<SIGN>(x[d], w);
    xa_put3<P, 1>(x, tid, buf);
    xa_get3<P, 2>(x, w, tid, buf, twb);
#pragma unroll
    for (int d = 0; d < 3; d++) xa_bfly<SIGN>(x[d], w);
    xa_put3<P, 2>(x, tid, buf);
    xa_get3<P, 3>(x, w, tid, buf, twb);
#pragma unroll
    for (int d = 0; d < 3; d++) xa_bfly<SIGN>(x[d], w);
}

// The exchange buffer of a line transform with LDS of its own (allocated for
// a kernel only when the kernel calls this).
template <int LS>
__device__ __forceinline__ double* xd_exchange() {
    __shared__ double ex[2 * LS];
    return ex;
}

// E with its inputs streamed into LDS (FP64, one wave per x-line, a
// persistent grid of one-wave blocks looping over the lines; no body force).
// The same arithmetic as k_ns_fourier_x with EMIT.  A line's three products
// and its V1 and N0 arrive in LDS by direct global->LDS loads issued a line
// ahead: the next line's products once this line's product transforms have
// read theirs, the next line's V1 and N0 once this line's update has read
// its own -- so the loads are in flight during the update, the emitted
// inverse transforms and the next product transforms.  All reads of the
// landed data are inline-asm LDS reads that wait for themselves, the twiddles
// come from an LDS copy, and the FFT exchanges synchronise with SyncLds: the
// only vector-memory operations left are the loads into LDS and the stores,
// and each wait counts them explicitly (vmcnt counts them in issue order).
// SHARE: the exchanges of the product transforms use the product slots (read
// into registers first) and those of the emitted inverse transforms the N0
// slots, whose next loads are then issued after the emitted transforms; the
// exchanges are the inline-asm ones of line_fft_xa (swizzled exchange index,
// stage-major twiddles).  The block's LDS shrinks
// from 45.3 KB to 40 KB, so four one-wave blocks fit on a compute unit (one
// per SIMD) instead of three.  G3 (with SHARE): the three product transforms,
// and the emitted ones in two groups of three, run together stage by stage
// (line_fft_xa3).
// YS (with SHARE and G3, Crank-Nicolson, single box): the merged AB2 history.
// The update needs V1 and N0 only through Y = (V1 + lap1) - K/2 N0, so the N0
// fields hold Y instead, V1 is neither read nor stored, and the kernel reads
// 6 and writes 9 field passes instead of 9 and 12:
//     S = Y + K/2 (3 N1)                  (= (V1 + lap1) + F)
//     V = (S - grad G) / (1 + nu K/2 |k|^2),   G = lap^-1 div S
//     Y' = (V + nu K/2 lap V) - K/2 N1    (the next step's Y)
// The velocity is then held only as its unnormalised x-inverse in the first
// three W fields (the next step's input); the host recovers V1 from them by
// a forward x pass when it needs it.  The same values as the (V1, N0) form to
// rounding (a different association of the same sums).
// E5 (with YS): the emitted fields are the x-inverses of V and of i kx Vy,
// i kx Vz -- five instead of six; the y pass forms the vorticity from them
// on load (k_ns_y_curl), since omega's remaining terms need only ky and kz.
// REV: the lines in descending order (the top planes first, which the
// preceding y pass wrote last and the last-level cache may still hold)
// LAUX: cache-policy bits of the loads into LDS (2: nontemporal, so the
// inputs, read once, do not displace in the last-level cache what the
// kernel writes for the next one)
template <class P, bool NT = true, bool SHARE = false, bool G3 = false, bool YS = false, bool E5 = false,
          bool REV = false, int LAUX = 0>
__global__ __launch_bounds__(64) void k_ns_fourier_xd(double2* __restrict__ W, double2* __restrict__ V,
                                                      double2* __restrict__ N0, const double2* __restrict__ twg,
                                                      int ny, int nz, int nzp, long F,
                                                      const StepCoef<double>* __restrict__ cfs,
                                                      double* __restrict__ divpart, long boff, int divstride) {
    using T = double;
    constexpr int N = P::N;
    constexpr int LS = LdsLine<N>::LS;
    constexpr int E = P::E;
    constexpr int R0 = P::radix(0);
    constexpr int Q = N / R0;
    constexpr int PIECES = N / 64;  // 64-value pieces of a line (one load into LDS each)
    static_assert(P::TPL == 64 && R0 == 4 && P::nb(0) == 1 && P::full(0) && P::radix(P::S - 1) == 4 &&
                      P::nb(P::S - 1) == 1 && N % 64 == 0 && 3 * Q * 16 < 65536,
                  "one wave per line, one radix-4 butterfly per lane in the first and last stage");
    static_assert(!SHARE || sm_size<P>() <= N, "stage-major twiddles fit the table");
    static_assert(!YS || (SHARE && G3), "the merged history runs the grouped, shared-slot kernel");
    static_assert(!E5 || YS, "five emitted fields with the merged history only");
    // vector-memory operations a line issues after its products' loads (the
    // next V1, the 6 + 6 field stores and, when SHARE, the next N0; with YS
    // the 3 + 6 (E5: 3 + 5) stores and the next Y), capped at what vmcnt encodes
    constexpr int AFTER_X = YS ? 3 * PIECES + (E5 ? 8 : 9) * E : (SHARE ? 6 * PIECES : 3 * PIECES) + 12 * E;
    constexpr int TOP_WAIT = SHARE ? (AFTER_X < 63 ? AFTER_X : 63) : 6 * PIECES;
    const StepCoef<T> cf = cfs[blockIdx.y];
    W += blockIdx.y * boff;
    V += blockIdx.y * boff;
    N0 += blockIdx.y * boff;
    divpart += (long)blockIdx.y * divstride;
    __shared__ double2 xs[3][N];                        // the products of the line
    __shared__ double2 v1l[YS ? 1 : 3][YS ? 1 : N];     // its V1 (not read with YS)
    __shared__ double2 n0l[3][N];                       // its N0 (YS: its Y)
    __shared__ double2 tw[N];
    T* sre = nullptr;  // the exchange of the transforms when it has LDS of its own
    if constexpr (!SHARE) sre = xd_exchange<LS>();
    T* const sim = SHARE ? nullptr : sre + LS;
    const int tid = threadIdx.x;
    if constexpr (SHARE)
        sm_fill<P, 1, T>(tw, twg, tid, 64);  // stage-major, for line_fft_xa
    else
        for (int t = tid; t < N; t += 64) tw[t] = twg[t];
    __syncthreads();
    const long nlines = (long)ny * nzp;
    auto idx = [&](int pos) { return pos + (pos >> 4); };
    auto fetch = [&](double2* slots, const double2* src, int nf, long line) {
        for (int f = 0; f < nf; f++)
#pragma unroll
            for (int k = 0; k < PIECES; k++)
                __builtin_amdgcn_global_load_lds(
                    (const __attribute__((address_space(1))) void*)(src + f * F + line * N + k * 64 + tid),
                    (__attribute__((address_space(3))) void*)(slots + f * N + k * 64), 16, 0, LAUX);
    };
    typedef double d2v __attribute__((ext_vector_type(2)));
    // this lane's first-stage positions tid + r Q of a landed field
    auto read4 = [&](const double2* slot, C2<T>* x) {
        const unsigned a = lds_offset(slot + tid);
        d2v v0, v1, v2, v3;
        asm volatile(
            "ds_read_b128 %0, %4\n\t"
            "ds_read_b128 %1, %4 offset:%5\n\t"
            "ds_read_b128 %2, %4 offset:%6\n\t"
            "ds_read_b128 %3, %4 offset:%7\n\t"
            "s_waitcnt lgkmcnt(0)"
            : "=&v"(v0), "=&v"(v1), "=&v"(v2), "=&v"(v3)
            : "v"(a), "n"(Q * 16), "n"(2 * Q * 16), "n"(3 * Q * 16)
            : "memory");
        x[0] = mk<T>(v0.x, v0.y);
        x[1] = mk<T>(v1.x, v1.y);
        x[2] = mk<T>(v2.x, v2.y);
        x[3] = mk<T>(v3.x, v3.y);
    };
    double dsum = 0;
    // the line a block's k-th trip handles (the trips themselves ascend)
    auto at = [&](long t) { return REV ? nlines - 1 - t : t; };
    long line = blockIdx.x;
    if (line < nlines) {
        fetch(&xs[0][0], W, 3, at(line));
        if constexpr (!YS) fetch(&v1l[0][0], V, 3, at(line));
        fetch(&n0l[0][0], N0, 3, at(line));
        // the first line's products (TOP_WAIT may exceed what follows them here)
        if constexpr (SHARE) asm volatile("s_waitcnt vmcnt(%0)" ::"n"((YS ? 3 : 6) * PIECES) : "memory");
    }
    for (; line < nlines; line += gridDim.x) {
        const long nxt = line + gridDim.x;
        const bool more = nxt < nlines;
        const long pl = at(line);
        const int jj = (int)(pl % ny), kk = (int)(pl / ny);
        const T k1 = freq<T>(jj, ny), k2 = freq<T>(kk, nz);
        const bool zline = (jj == ny / 2) || (kk == nz / 2);
        const double dw = (nzp < nz && kk > 0) ? 2.0 : 1.0;
        const long off = pl * N;
        // products of this line landed: issued after them are this line's V1 and
        // N0 and, from the second line on, the previous line's stores (and, when
        // SHARE, only V1 ahead of the stores and N0 after them)
        asm volatile("s_waitcnt vmcnt(%0)" ::"n"(TOP_WAIT) : "memory");
        C2<T> n1[3][E], x[E];
        if constexpr (SHARE) {
#pragma unroll
            for (int d = 0; d < 3; d++) read4(&xs[d][0], n1[d]);
            // the exchanges reuse the product slots; their last reads have
            // completed before the next products are loaded there
            if constexpr (G3) {
                line_fft_xa3<P, -1>(n1, tid, lds_offset(&xs[0][0]), tw);
            } else {
#pragma unroll
                for (int d = 0; d < 3; d++) line_fft_xa<P, -1>(n1[d], tid, lds_offset(&xs[0][0]), tw);
            }
        } else {
#pragma unroll
            for (int d = 0; d < 3; d++) {
                read4(&xs[d][0], x);
                if (d > 0) SyncLds::run();
                line_fft<P, -1, T, decltype(idx), SyncLds>(x, tid, sre, sim, idx, tw);
#pragma unroll
                for (int s2 = 0; s2 < E; s2++) n1[d][s2] = x[s2];
            }
        }
        if (more) fetch(&xs[0][0], W, 3, at(nxt));
        // V1 and N0 of this line: issued after them, the previous line's stores
        // and the next line's products (none of either on a block's only line)
        if (more)
            asm volatile("s_waitcnt vmcnt(%0)" ::"n"(3 * PIECES) : "memory");
        else
            asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
        C2<T> v1s[YS ? 1 : 3][E], n0s[3][E], vn[3][E];
#pragma unroll
        for (int d = 0; d < 3; d++) {
            if constexpr (!YS) read4(&v1l[d][0], v1s[d]);
            read4(&n0l[d][0], n0s[d]);
        }
        if (more) {
            if constexpr (!YS) fetch(&v1l[0][0], V, 3, at(nxt));
            if constexpr (!SHARE) fetch(&n0l[0][0], N0, 3, at(nxt));
        }
        if constexpr (YS) {
            for_out<P>(tid, [&](int s, int pos) {
                const T k0 = freq<T>(pos, N);
                const bool z = zline || pos == N / 2 ||
                               (cf.dealias == 1 && (3 * fabs(k0) > N || 3 * fabs(k1) > ny || 3 * fabs(k2) > nz));
                const T sig =
                    cf.dealias == 2 ? expfilter<T>(k0, N) * expfilter<T>(k1, ny) * expfilter<T>(k2, nz) : T(1);
                C2<T> nn[3], S[3];
#pragma unroll
                for (int d = 0; d < 3; d++) {
                    nn[d] = z ? mk<T>(T(0), T(0)) : n1[d][s];
                    if (cf.dealias == 2) nn[d] = rmul<T>(sig, nn[d]);
                    // S = Y + K/2*(3.*cross1)  (= (V1 + lap1) + F)
                    S[d] = cadd<T>(n0s[d][s], rmul<T>(cf.Kh, rmul<T>(T(3), nn[d])));
                }
                const T q = -(k0 * k0 + k1 * k1 + k2 * k2);
                // G = div(S); G.laplacian_inverse() (a=0, b=1)
                C2<T> G = mk<T>(T(0), T(0));
                G = cadd<T>(G, imul<T>(k0, S[0]));
                G = cadd<T>(G, imul<T>(k1, S[1]));
                G = cadd<T>(G, imul<T>(k2, S[2]));
                T sum = 0;
                sum += k0 * k0;
                sum += k1 * k1;
                sum += k2 * k2;
                const T rg = sum != 0 ? T(1) / (T(0) - T(1) * sum) : T(1);
                G.x *= rg;
                G.y *= rg;
                const double rv = 1.0 / (1.0 - (-cf.c1d) * (double)sum);  // 1 where sum == 0
                const T kk3[3] = {k0, k1, k2};
#pragma unroll
                for (int d = 0; d < 3; d++) {
                    // V = Vhalf + lap1 = S - G.gradient(), then the CN divide (FP64 factor)
                    const C2<T> vv = csub<T>(S[d], imul<T>(kk3[d], G));
                    vn[d][s] = mk<T>(T((double)vv.x * rv), T((double)vv.y * rv));
                }
                C2<T> dv = mk<T>(T(0), T(0));
                dv = cadd<T>(dv, imul<T>(k0, vn[0][s]));
                dv = cadd<T>(dv, imul<T>(k1, vn[1][s]));
                dv = cadd<T>(dv, imul<T>(k2, vn[2][s]));
                dsum += dw * sqrt((double)dv.x * (double)dv.x + (double)dv.y * (double)dv.y);
#pragma unroll
                for (int d = 0; d < 3; d++) {
                    // the next step's Y = (V + lap V) - K/2 cross1
                    const C2<T> yn = csub<T>(cadd<T>(vn[d][s], rmul<T>(cf.c1, rmul<T>(q, vn[d][s]))),
                                             rmul<T>(cf.Kh, nn[d]));
                    st_line<NT, T>(N0 + d * F + off + pos, yn);
                }
            });
        } else
        for_out<P>(tid, [&](int s, int pos) {
            const T k0 = freq<T>(pos, N);
            const bool z = zline || pos == N / 2 ||
                           (cf.dealias == 1 && (3 * fabs(k0) > N || 3 * fabs(k1) > ny || 3 * fabs(k2) > nz));
            const T sig = cf.dealias == 2 ? expfilter<T>(k0, N) * expfilter<T>(k1, ny) * expfilter<T>(k2, nz) : T(1);
            C2<T> nn[3], v1[3], f[3], L[3], S[3];
#pragma unroll
            for (int d = 0; d < 3; d++) {
                nn[d] = z ? mk<T>(T(0), T(0)) : n1[d][s];
                if (cf.dealias == 2) nn[d] = rmul<T>(sig, nn[d]);
                v1[d] = v1s[d][s];
                const C2<T> n0 = n0s[d][s];
                // F = K/2*(3.*cross1 - cross0)
                f[d] = rmul<T>(cf.Kh, csub<T>(rmul<T>(T(3), nn[d]), n0));
            }
            // lap1 = nu*K/2*V1.laplacian() = c1 * (-(k.k) V1)
            const T q = -(k0 * k0 + k1 * k1 + k2 * k2);
#pragma unroll
            for (int d = 0; d < 3; d++) {
                L[d] = rmul<T>(cf.c1, rmul<T>(q, v1[d]));
                S[d] = cadd<T>(cadd<T>(v1[d], L[d]), f[d]);  // (V1 + lap1) + F
            }
            // G = div(S); G.laplacian_inverse() (a=0, b=1)
            C2<T> G = mk<T>(T(0), T(0));
            G = cadd<T>(G, imul<T>(k0, S[0]));
            G = cadd<T>(G, imul<T>(k1, S[1]));
            G = cadd<T>(G, imul<T>(k2, S[2]));
            T sum = 0;
            sum += k0 * k0;
            sum += k1 * k1;
            sum += k2 * k2;
            // one reciprocal per point for each factor, shared by the components
            // (within an ulp of dividing each component)
            const T rg = sum != 0 ? T(1) / (T(0) - T(1) * sum) : T(1);
            G.x *= rg;
            G.y *= rg;
            const double rv = 1.0 / (1.0 - (-cf.c1d) * (double)sum);  // 1 where sum == 0
            const T kk3[3] = {k0, k1, k2};
#pragma unroll
            for (int d = 0; d < 3; d++) {
                // Vhalf = V1 + F - G.gradient();  V = Vhalf + lap1 (exec/ns2.cpp:185: V = Vhalf)
                C2<T> vh = csub<T>(cadd<T>(v1[d], f[d]), imul<T>(kk3[d], G));
                C2<T> vv = cf.be ? vh : cadd<T>(vh, L[d]);
                // V[d].laplacian_inverse(1, -nu*K/2): divide by 1 + (nu K/2)|k|^2 (FP64 factor)
                vv = mk<T>(T((double)vv.x * rv), T((double)vv.y * rv));
                vn[d][s] = vv;
            }
            C2<T> dv = mk<T>(T(0), T(0));
            dv = cadd<T>(dv, imul<T>(k0, vn[0][s]));
            dv = cadd<T>(dv, imul<T>(k1, vn[1][s]));
            dv = cadd<T>(dv, imul<T>(k2, vn[2][s]));
            dsum += dw * sqrt((double)dv.x * (double)dv.x + (double)dv.y * (double)dv.y);  // |div V| (no overflow here)
#pragma unroll
            for (int d = 0; d < 3; d++) {
                st_line<NT, T>(V + d * F + off + pos, vn[d][s]);
                st_line<NT, T>(N0 + d * F + off + pos, nn[d]);
            }
        });
        // the next step's kernel A: inverse x of V (g = 0..2) and of i k x V (g = 3..5)
        if constexpr (G3) {
            static_assert(SHARE && E == 4, "grouped transforms in the N0 slots");
            C2<T> x3[3][E];
#pragma unroll
            for (int h = 0; h < 2; h++) {
                for_out<P>(tid, [&](int s, int pos) {
                    const T k0 = freq<T>(pos, N);
                    if (h == 0) {
#pragma unroll
                        for (int d = 0; d < 3; d++) x3[d][s] = vn[d][s];
                    } else if constexpr (E5) {
                        x3[0][s] = imul<T>(k0, vn[1][s]);  // d/dx Vy
                        x3[1][s] = imul<T>(k0, vn[2][s]);  // d/dx Vz
                        x3[2][s] = mk<T>(T(0), T(0));      // (transformed, not stored)
                    } else {
                        x3[0][s] = csub<T>(imul<T>(k1, vn[2][s]), imul<T>(k2, vn[1][s]));
                        x3[1][s] = csub<T>(imul<T>(k2, vn[0][s]), imul<T>(k0, vn[2][s]));
                        x3[2][s] = csub<T>(imul<T>(k0, vn[1][s]), imul<T>(k1, vn[0][s]));
                    }
                });
                line_fft_xa3<P, 1>(x3, tid, lds_offset(&n0l[0][0]), tw);  // in the N0 slots, read above
#pragma unroll
                for (int d = 0; d < ((E5 && h == 1) ? 2 : 3); d++)
                    for_out<P>(tid, [&](int s, int pos) { st_line<NT, T>(W + (3 * h + d) * F + off + pos, x3[d][s]); });
            }
        }
#pragma unroll
        for (int g = 0; g < (G3 ? 0 : 6); g++) {
            for_out<P>(tid, [&](int s, int pos) {
                const T k0 = freq<T>(pos, N);
                if (g < 3)
                    x[s] = vn[g][s];
                else if (g == 3)
                    x[s] = csub<T>(imul<T>(k1, vn[2][s]), imul<T>(k2, vn[1][s]));
                else if (g == 4)
                    x[s] = csub<T>(imul<T>(k2, vn[0][s]), imul<T>(k0, vn[2][s]));
                else
                    x[s] = csub<T>(imul<T>(k0, vn[1][s]), imul<T>(k1, vn[0][s]));
            });
            if constexpr (SHARE) {
                line_fft_xa<P, 1>(x, tid, lds_offset(&n0l[0][0]), tw);  // in the N0 slots, read above
            } else {
                SyncLds::run();
                line_fft<P, 1, T, decltype(idx), SyncLds>(x, tid, sre, sim, idx, tw);
            }
            for_out<P>(tid, [&](int s, int pos) { st_line<NT, T>(W + g * F + off + pos, x[s]); });
        }
        if constexpr (SHARE) {
            if (more) fetch(&n0l[0][0], N0, 3, at(nxt));
        } else {
            SyncLds::run();
        }
    }
    // this block's sum of |div V| over its lines (FP64; lanes in a fixed order)
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) dsum += __shfl_xor(dsum, o, 64);
    if (tid == 0) divpart[blockIdx.x] = dsum;
}

// ---- FP32 kernel E with its inputs streamed into LDS
// The exchanges of line_fft_xa3 with 8-byte complex values (ds_write_b64 /
// ds_read_b64 at the same swizzled index), in inline asm for the same reason:
// the compiler then inserts no wait for the loads into LDS in flight.
typedef float xf_f2 __attribute__((ext_vector_type(2)));

template <class P, int s>
__device__ __forceinline__ void xf_put3(const C2<float> (&x)[3][4], int tid, unsigned buf) {
    constexpr int R = P::radix(s), Ns = P::ns(s);
    const int base = (tid / Ns) * Ns * R + (tid % Ns);
#pragma unroll
    for (int r = 0; r < R; r++) {
        const unsigned a = buf + 8u * (unsigned)SwzIdx()(base + r * Ns);
#pragma unroll
        for (int d = 0; d < 3; d++) {
            const xf_f2 v = {x[d][r].x, x[d][r].y};
            asm volatile("ds_write_b64 %0, %1 offset:%2" ::"v"(a), "v"(v), "n"(d * 8 * P::N) : "memory");
        }
    }
}

// the values tid + r N/4 of each line after stage s - 1 and this lane's three
// twiddles of stage s, in one asm block that waits for them
template <class P, int s>
__device__ __forceinline__ void xf_get3(C2<float> (&x)[3][4], C2<float>* w, int tid, unsigned buf,
                                        unsigned twb) {
    constexpr int Q = P::N / 4;
    constexpr int D = 8 * Q;
    constexpr int Ns = P::ns(s);
    static_assert(Q % 64 == 0 && 2 * 8 * P::N + 3 * D < 65536 && 16 * Ns < 65536, "constant offsets");
    const unsigned a0 = buf + 8u * (unsigned)SwzIdx()(tid);
    const unsigned a1 = a0 + 8u * P::N, a2 = a1 + 8u * P::N;
    const unsigned t = twb + 8u * (unsigned)(sm_off<P, s>() + tid % Ns);
    xf_f2 v[3][4], w1, w2, w3;
    asm volatile(
        "ds_read_b64 %0, %15\n\t"
        "ds_read_b64 %1, %15 offset:%19\n\t"
        "ds_read_b64 %2, %15 offset:%20\n\t"
        "ds_read_b64 %3, %15 offset:%21\n\t"
        "ds_read_b64 %4, %16\n\t"
        "ds_read_b64 %5, %16 offset:%19\n\t"
        "ds_read_b64 %6, %16 offset:%20\n\t"
        "ds_read_b64 %7, %16 offset:%21\n\t"
        "ds_read_b64 %8, %17\n\t"
        "ds_read_b64 %9, %17 offset:%19\n\t"
        "ds_read_b64 %10, %17 offset:%20\n\t"
        "ds_read_b64 %11, %17 offset:%21\n\t"
        "ds_read_b64 %12, %18\n\t"
        "ds_read_b64 %13, %18 offset:%22\n\t"
        "ds_read_b64 %14, %18 offset:%23\n\t"
        "s_waitcnt lgkmcnt(0)"
        : "=&v"(v[0][0]), "=&v"(v[0][1]), "=&v"(v[0][2]), "=&v"(v[0][3]), "=&v"(v[1][0]), "=&v"(v[1][1]),
          "=&v"(v[1][2]), "=&v"(v[1][3]), "=&v"(v[2][0]), "=&v"(v[2][1]), "=&v"(v[2][2]), "=&v"(v[2][3]),
          "=&v"(w1), "=&v"(w2), "=&v"(w3)
        : "v"(a0), "v"(a1), "v"(a2), "v"(t), "n"(D), "n"(2 * D), "n"(3 * D), "n"(8 * Ns), "n"(16 * Ns)
        : "memory");
#pragma unroll
    for (int d = 0; d < 3; d++)
#pragma unroll
        for (int r = 0; r < 4; r++) x[d][r] = mk<float>(v[d][r].x, v[d][r].y);
    w[0] = mk<float>(w1.x, w1.y);
    w[1] = mk<float>(w2.x, w2.y);
    w[2] = mk<float>(w3.x, w3.y);
}

template <int SIGN>
__device__ __forceinline__ void xf_bfly(C2<float>* x, const C2<float>* w) {
#pragma unroll
    for (int r = 1; r < 4; r++) {
        C2<float> t = w[r - 1];
        if constexpr (SIGN > 0) t.y = -t.y;
        x[r] = cmul<float>(x[r], t);
    }
    Dft<float, 4, SIGN>::run(x);
}

// three 256-point lines of one wave, stage by stage (line_fft_xa3 in FP32)
template <class P, int SIGN>
__device__ __forceinline__ void line_fft_xf3(C2<float> (&x)[3][4], int tid, unsigned buf,
                                             const C2<float>* __restrict__ tws) {
    static_assert(P::TPL == 64 && P::S == 4 && P::radix(0) == 4 && P::radix(1) == 4 && P::radix(2) == 4 &&
                      P::radix(3) == 4 && P::N == 256 && P::ns(0) == 1,
                  "four radix-4 stages, one butterfly per lane");
    const unsigned twb = lds_offset(tws);
    C2<float> w[3];
#pragma unroll
    for (int d = 0; d < 3; d++) stage_bfly<P, 0, SIGN, float>(x[d], tid, tws);  // no twiddles
    xf_put3<P, 0>(x, tid, buf);
    xf_get3<P, 1>(x, w, tid, buf, twb);
#pragma unroll
    for (int d = 0; d < 3; d++) xf_bfly<SIGN>(x[d], w);
    xf_put3<P, 1>(x, tid, buf);
    xf_get3<P, 2>(x, w, tid, buf, twb);
#pragma unroll
    for (int d = 0; d < 3; d++) xf_bfly<SIGN>(x[d], w);
    xf_put3<P, 2>(x, tid, buf);
    xf_get3<P, 3>(x, w, tid, buf, twb);
#pragma unroll
    for (int d = 0; d < 3; d++) xf_bfly<SIGN>(x[d], w);
}

// The same for the 512-point plan P512 (three radix-8 stages, one butterfly
// per lane): each line's eight reads in an asm block of their own that waits
// for them, the stage's seven twiddles in another.
template <class P, int s>
__device__ __forceinline__ void xf8_put3(const C2<float> (&x)[3][8], int tid, unsigned buf) {
    constexpr int R = P::radix(s), Ns = P::ns(s);
    static_assert(R == 8 && 2 * 8 * P::N < 65536, "radix-8 stages, constant offsets");
    const int base = (tid / Ns) * Ns * R + (tid % Ns);
#pragma unroll
    for (int r = 0; r < R; r++) {
        const unsigned a = buf + 8u * (unsigned)SwzIdx()(base + r * Ns);
#pragma unroll
        for (int d = 0; d < 3; d++) {
            const xf_f2 v = {x[d][r].x, x[d][r].y};
            asm volatile("ds_write_b64 %0, %1 offset:%2" ::"v"(a), "v"(v), "n"(d * 8 * P::N) : "memory");
        }
    }
}

template <class P, int s>
__device__ __forceinline__ void xf8_get3(C2<float> (&x)[3][8], C2<float>* w, int tid, unsigned buf,
                                         unsigned twb) {
    constexpr int Q = P::N / 8;
    constexpr int D = 8 * Q;
    constexpr int Ns = P::ns(s);
    static_assert(Q % 64 == 0 && 2 * 8 * P::N + 7 * D < 65536 && 48 * Ns < 65536, "constant offsets");
    const unsigned a0 = buf + 8u * (unsigned)SwzIdx()(tid);
#pragma unroll
    for (int d = 0; d < 3; d++) {
        const unsigned a = a0 + (unsigned)d * 8u * P::N;
        xf_f2 v[8];
        asm volatile(
            "ds_read_b64 %0, %8\n\t"
            "ds_read_b64 %1, %8 offset:%9\n\t"
            "ds_read_b64 %2, %8 offset:%10\n\t"
            "ds_read_b64 %3, %8 offset:%11\n\t"
            "ds_read_b64 %4, %8 offset:%12\n\t"
            "ds_read_b64 %5, %8 offset:%13\n\t"
            "ds_read_b64 %6, %8 offset:%14\n\t"
            "ds_read_b64 %7, %8 offset:%15\n\t"
            "s_waitcnt lgkmcnt(0)"
            : "=&v"(v[0]), "=&v"(v[1]), "=&v"(v[2]), "=&v"(v[3]), "=&v"(v[4]), "=&v"(v[5]), "=&v"(v[6]),
              "=&v"(v[7])
            : "v"(a), "n"(D), "n"(2 * D), "n"(3 * D), "n"(4 * D), "n"(5 * D), "n"(6 * D), "n"(7 * D)
            : "memory");
#pragma unroll
        for (int r = 0; r < 8; r++) x[d][r] = mk<float>(v[r].x, v[r].y);
    }
    const unsigned t = twb + 8u * (unsigned)(sm_off<P, s>() + tid % Ns);
    xf_f2 u[7];
    asm volatile(
        "ds_read_b64 %0, %7\n\t"
        "ds_read_b64 %1, %7 offset:%8\n\t"
        "ds_read_b64 %2, %7 offset:%9\n\t"
        "ds_read_b64 %3, %7 offset:%10\n\t"
        "ds_read_b64 %4, %7 offset:%11\n\t"
        "ds_read_b64 %5, %7 offset:%12\n\t"
        "ds_read_b64 %6, %7 offset:%13\n\t"
        "s_waitcnt lgkmcnt(0)"
        : "=&v"(u[0]), "=&v"(u[1]), "=&v"(u[2]), "=&v"(u[3]), "=&v"(u[4]), "=&v"(u[5]), "=&v"(u[6])
        : "v"(t), "n"(8 * Ns), "n"(16 * Ns), "n"(24 * Ns), "n"(32 * Ns), "n"(40 * Ns), "n"(48 * Ns)
        : "memory");
#pragma unroll
    for (int r = 0; r < 7; r++) w[r] = mk<float>(u[r].x, u[r].y);
}

template <int SIGN>
__device__ __forceinline__ void xf8_bfly(C2<float>* x, const C2<float>* w) {
#pragma unroll
    for (int r = 1; r < 8; r++) {
        C2<float> t = w[r - 1];
        if constexpr (SIGN > 0) t.y = -t.y;
        x[r] = cmul<float>(x[r], t);
    }
    Dft<float, 8, SIGN>::run(x);
}

template <class P, int SIGN>
__device__ __forceinline__ void line_fft_xf3(C2<float> (&x)[3][8], int tid, unsigned buf,
                                             const C2<float>* __restrict__ tws) {
    static_assert(P::TPL == 64 && P::S == 3 && P::radix(0) == 8 && P::radix(1) == 8 && P::radix(2) == 8 &&
                      P::N == 512 && P::ns(0) == 1,
                  "three radix-8 stages, one butterfly per lane");
    const unsigned twb = lds_offset(tws);
    C2<float> w[7];
#pragma unroll
    for (int d = 0; d < 3; d++) stage_bfly<P, 0, SIGN, float>(x[d], tid, tws);  // no twiddles
    xf8_put3<P, 0>(x, tid, buf);
    xf8_get3<P, 1>(x, w, tid, buf, twb);
#pragma unroll
    for (int d = 0; d < 3; d++) xf8_bfly<SIGN>(x[d], w);
    xf8_put3<P, 1>(x, tid, buf);
    xf8_get3<P, 2>(x, w, tid, buf, twb);
#pragma unroll
    for (int d = 0; d < 3; d++) xf8_bfly<SIGN>(x[d], w);
}

// Kernel E of the merged AB2 history in FP32, its inputs streamed into LDS:
// the arithmetic of k_ns_fourier_x with YS (and E5) in the structure of
// k_ns_fourier_xd with SHARE, G3 and YS -- one wave per x line, a persistent
// grid of one-wave blocks; a line's products and its Y arrive in LDS by
// direct global->LDS loads issued a line ahead (16 bytes, two values, per
// lane), the product transforms exchange in the product slots, the emitted
// inverse transforms in the Y slots, and every LDS access in the loop is
// inline asm that waits for itself.  Single box, no body force.
template <class P, bool NT, bool E5, bool REV, int LAUX>
__global__ __launch_bounds__(64) void k_ns_fourier_xs(
    float2* __restrict__ W, float2* __restrict__ N0,
                                                      const float2* __restrict__ twg, int ny, int nz, int nzp,
                                                      long F, const StepCoef<float>* __restrict__ cfs,
                                                      double* __restrict__ divpart) {
    using T = float;
    constexpr int N = P::N;
    constexpr int E = P::E;
    constexpr int Q = N / E;
    constexpr int PIECES = N / 128;  // 128-value pieces of a line (one load into LDS each)
    static_assert(P::TPL == 64 && (E == 4 || E == 8) && P::radix(0) == E && N % 128 == 0 && sm_size<P>() <= N,
                  "one wave per line, one radix-4 or radix-8 butterfly per lane; stage-major twiddles fit the table");
    // vector-memory operations a line issues after its products' loads: the
    // 3 + 6 (E5: 3 + 5) field stores and the next line's Y
    constexpr int AFTER_X = 3 * PIECES + (E5 ? 8 : 9) * E;
    constexpr int TOP_WAIT = AFTER_X < 63 ? AFTER_X : 63;
    const StepCoef<T> cf = cfs[0];
    __shared__ float2 xs[3][N];   // the products of the line
    __shared__ float2 n0l[3][N];  // its Y
    __shared__ float2 tw[N];
    const int tid = threadIdx.x;
    sm_fill<P, 1, T>(tw, twg, tid, 64);  // stage-major
    __syncthreads();
    const long nlines = (long)ny * nzp;
    auto fetch = [&](float2* slots, const float2* src, long line) {
#pragma unroll
        for (int f = 0; f < 3; f++)
#pragma unroll
            for (int k = 0; k < PIECES; k++)
                __builtin_amdgcn_global_load_lds(
                    (const __attribute__((address_space(1))) void*)(src + f * F + line * N + k * 128 + 2 * tid),
                    (__attribute__((address_space(3))) void*)(slots + f * N + k * 128), 16, 0, LAUX);
    };
    // this lane's first-stage positions tid + r Q of a landed field (in
    // groups of four reads, each asm block waiting for its own)
    auto read4 = [&](const float2* slot, C2<T>* x) {
#pragma unroll
        for (int g = 0; g < E / 4; g++) {
            const unsigned a = lds_offset(slot + tid + 4 * g * Q);
            xf_f2 v0, v1, v2, v3;
            asm volatile(
                "ds_read_b64 %0, %4\n\t"
                "ds_read_b64 %1, %4 offset:%5\n\t"
                "ds_read_b64 %2, %4 offset:%6\n\t"
                "ds_read_b64 %3, %4 offset:%7\n\t"
                "s_waitcnt lgkmcnt(0)"
                : "=&v"(v0), "=&v"(v1), "=&v"(v2), "=&v"(v3)
                : "v"(a), "n"(Q * 8), "n"(2 * Q * 8), "n"(3 * Q * 8)
                : "memory");
            x[4 * g + 0] = mk<T>(v0.x, v0.y);
            x[4 * g + 1] = mk<T>(v1.x, v1.y);
            x[4 * g + 2] = mk<T>(v2.x, v2.y);
            x[4 * g + 3] = mk<T>(v3.x, v3.y);
        }
    };
    double dsum = 0;
    // the line a block's k-th trip handles (the trips themselves ascend)
    auto at = [&](long t) { return REV ? nlines - 1 - t : t; };
    long line = blockIdx.x;
    if (line < nlines) {
        fetch(&xs[0][0], W, at(line));
        fetch(&n0l[0][0], N0, at(line));
        asm volatile("s_waitcnt vmcnt(%0)" ::"n"(3 * PIECES) : "memory");  // the first line's products
    }
    for (; line < nlines; line += gridDim.x) {
        const long nxt = line + gridDim.x;
        const bool more = nxt < nlines;
        const long pl = at(line);
        const int jj = (int)(pl % ny), kk = (int)(pl / ny);
        const T k1 = freq<T>(jj, ny), k2 = freq<T>(kk, nz);
        const bool zline = (jj == ny / 2) || (kk == nz / 2);
        const double dw = (nzp < nz && kk > 0) ? 2.0 : 1.0;
        const long off = pl * N;
        // products of this line landed: issued after them are this line's Y
        // and, from the second line on, the previous line's stores
        asm volatile("s_waitcnt vmcnt(%0)" ::"n"(TOP_WAIT) : "memory");
        C2<T> n1[3][E];
#pragma unroll
        for (int d = 0; d < 3; d++) read4(&xs[d][0], n1[d]);
        // the exchanges reuse the product slots; their last reads have
        // completed before the next products are loaded there
        line_fft_xf3<P, -1>(n1, tid, lds_offset(&xs[0][0]), tw);
        if (more) fetch(&xs[0][0], W, at(nxt));
        // Y of this line: issued after it, the next line's products (none on a
        // block's last line)
        if (more)
            asm volatile("s_waitcnt vmcnt(%0)" ::"n"(3 * PIECES) : "memory");
        else
            asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
        C2<T> n0s[3][E], vn[3][E];
#pragma unroll
        for (int d = 0; d < 3; d++) read4(&n0l[d][0], n0s[d]);
        for_out<P>(tid, [&](int s, int pos) {
            const T k0 = freq<T>(pos, N);
            const bool z = zline || pos == N / 2 ||
                           (cf.dealias == 1 && (3 * fabs(k0) > N || 3 * fabs(k1) > ny || 3 * fabs(k2) > nz));
            const T sig = cf.dealias == 2 ? expfilter<T>(k0, N) * expfilter<T>(k1, ny) * expfilter<T>(k2, nz) : T(1);
            C2<T> nn[3], S[3];
#pragma unroll
            for (int d = 0; d < 3; d++) {
                nn[d] = z ? mk<T>(T(0), T(0)) : n1[d][s];
                if (cf.dealias == 2) nn[d] = rmul<T>(sig, nn[d]);
                // S = Y + K/2*(3.*cross1)  (= (V1 + lap1) + F)
                S[d] = cadd<T>(n0s[d][s], rmul<T>(cf.Kh, rmul<T>(T(3), nn[d])));
            }
            const T q = -(k0 * k0 + k1 * k1 + k2 * k2);
            // G = div(S); G.laplacian_inverse() (a=0, b=1)
            C2<T> G = mk<T>(T(0), T(0));
            G = cadd<T>(G, imul<T>(k0, S[0]));
            G = cadd<T>(G, imul<T>(k1, S[1]));
            G = cadd<T>(G, imul<T>(k2, S[2]));
            T sum = 0;
            sum += k0 * k0;
            sum += k1 * k1;
            sum += k2 * k2;
            const T rg = sum != 0 ? T(1) / (T(0) - T(1) * sum) : T(1);
            G.x *= rg;
            G.y *= rg;
            const double rv = 1.0 / (1.0 - (-cf.c1d) * (double)sum);  // 1 where sum == 0
            const T kk3[3] = {k0, k1, k2};
#pragma unroll
            for (int d = 0; d < 3; d++) {
                // V = Vhalf + lap1 = S - G.gradient(), then the CN divide (FP64 factor)
                const C2<T> vv = csub<T>(S[d], imul<T>(kk3[d], G));
                vn[d][s] = mk<T>(T((double)vv.x * rv), T((double)vv.y * rv));
            }
            C2<T> dv = mk<T>(T(0), T(0));
            dv = cadd<T>(dv, imul<T>(k0, vn[0][s]));
            dv = cadd<T>(dv, imul<T>(k1, vn[1][s]));
            dv = cadd<T>(dv, imul<T>(k2, vn[2][s]));
            dsum += dw * sqrt((double)dv.x * (double)dv.x + (double)dv.y * (double)dv.y);
#pragma unroll
            for (int d = 0; d < 3; d++) {
                // the next step's Y = (V + lap V) - K/2 cross1
                const C2<T> yn = csub<T>(cadd<T>(vn[d][s], rmul<T>(cf.c1, rmul<T>(q, vn[d][s]))),
                                         rmul<T>(cf.Kh, nn[d]));
                st_line<NT, T>(N0 + d * F + off + pos, yn);
            }
        });
        // the next step's kernel A: inverse x of V (group 0) and of i k x V
        // (group 1; E5: of i kx Vy, i kx Vz and a line of zeros, not stored)
        C2<T> x3[3][E];
#pragma unroll
        for (int h = 0; h < 2; h++) {
            for_out<P>(tid, [&](int s, int pos) {
                const T k0 = freq<T>(pos, N);
                if (h == 0) {
#pragma unroll
                    for (int d = 0; d < 3; d++) x3[d][s] = vn[d][s];
                } else if constexpr (E5) {
                    x3[0][s] = imul<T>(k0, vn[1][s]);  // d/dx Vy
                    x3[1][s] = imul<T>(k0, vn[2][s]);  // d/dx Vz
                    x3[2][s] = mk<T>(T(0), T(0));
                } else {
                    x3[0][s] = csub<T>(imul<T>(k1, vn[2][s]), imul<T>(k2, vn[1][s]));
                    x3[1][s] = csub<T>(imul<T>(k2, vn[0][s]), imul<T>(k0, vn[2][s]));
                    x3[2][s] = csub<T>(imul<T>(k0, vn[1][s]), imul<T>(k1, vn[0][s]));
                }
            });
            line_fft_xf3<P, 1>(x3, tid, lds_offset(&n0l[0][0]), tw);  // in the Y slots, read above
#pragma unroll
            for (int d = 0; d < ((E5 && h == 1) ? 2 : 3); d++)
                for_out<P>(tid, [&](int s, int pos) { st_line<NT, T>(W + (3 * h + d) * F + off + pos, x3[d][s]); });
        }
        if (more) fetch(&n0l[0][0], N0, at(nxt));
    }
    // this block's sum of |div V| over its lines (FP64; lanes in a fixed order)
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) dsum += __shfl_xor(dsum, o, 64);
    if (tid == 0) divpart[blockIdx.x] = dsum;
}

// ---- FP32 kernel E with wavefront exchanges
// A 256-point line held by one wave as 4 values per lane (lane l, register r:
// position l + 64 r), transformed with no LDS exchange: four radix-4 stages
// with the cross-lane swaps above between them.

template <typename T>
__device__ __forceinline__ void xw_twiddle(C2<T>* x, const C2<T>* w) {
#pragma unroll
    for (int r = 1; r < 4; r++) x[r] = cmul<T>(x[r], w[r - 1]);
}

// the base-4 digit reversal of a lane index (three digits)
__device__ __forceinline__ int xw_rev(int l) { return (l >> 4) + 4 * ((l >> 2) & 3) + 16 * (l & 3); }

// the table entries (of exp(-2 pi i k / 256)) of this lane's twiddles j = 1..3
// of forward stage st (0..2) and (conjugated) of inverse stage st
__device__ __forceinline__ int xw_fidx(int st, int j, int l) {
    return (st == 0 ? j * l : st == 1 ? 4 * j * (l & 15) : 16 * j * (l & 3)) & 255;
}
__device__ __forceinline__ int xw_iidx(int st, int j, int l) {
    return (st == 0 ? j * xw_rev(l) : st == 1 ? 4 * j * ((l >> 4) + 4 * ((l >> 2) & 3)) : 16 * j * (l >> 4)) & 255;
}

// FP32: this lane's twiddles held in registers (loaded once per kernel)
struct XwTw {
    C2<float> tf[3][3], ti[3][3];
    __device__ __forceinline__ void load(const float2* __restrict__ twg, int l) {
#pragma unroll
        for (int st = 0; st < 3; st++)
#pragma unroll
            for (int j = 1; j < 4; j++) {
                const float2 a = twg[xw_fidx(st, j, l)], b = twg[xw_iidx(st, j, l)];
                tf[st][j - 1] = mk<float>(a.x, a.y);
                ti[st][j - 1] = mk<float>(b.x, -b.y);
            }
    }
    __device__ __forceinline__ void fwd(int st, C2<float>* w, int) const {
#pragma unroll
        for (int j = 0; j < 3; j++) w[j] = tf[st][j];
    }
    __device__ __forceinline__ void inv(int st, C2<float>* w, int) const {
#pragma unroll
        for (int j = 0; j < 3; j++) w[j] = ti[st][j];
    }
};

// FP64: this lane's twiddles read at each stage from the natural table in
// LDS (inline asm that waits for itself)
struct XwTwLds {
    unsigned base;  // LDS address of the table
    __device__ __forceinline__ void get(int st, C2<double>* w, int l, bool inverse) const {
        typedef double d2 __attribute__((ext_vector_type(2)));
        d2 a, b, c;
        const unsigned i1 = base + 16u * (unsigned)(inverse ? xw_iidx(st, 1, l) : xw_fidx(st, 1, l));
        const unsigned i2 = base + 16u * (unsigned)(inverse ? xw_iidx(st, 2, l) : xw_fidx(st, 2, l));
        const unsigned i3 = base + 16u * (unsigned)(inverse ? xw_iidx(st, 3, l) : xw_fidx(st, 3, l));
        asm volatile(
            "ds_read_b128 %0, %3\n\t"
            "ds_read_b128 %1, %4\n\t"
            "ds_read_b128 %2, %5\n\t"
            "s_waitcnt lgkmcnt(0)"
            : "=&v"(a), "=&v"(b), "=&v"(c)
            : "v"(i1), "v"(i2), "v"(i3)
            : "memory");
        const double sg = inverse ? -1.0 : 1.0;
        w[0] = mk<double>(a.x, sg * a.y);
        w[1] = mk<double>(b.x, sg * b.y);
        w[2] = mk<double>(c.x, sg * c.y);
    }
    __device__ __forceinline__ void fwd(int st, C2<double>* w, int l) const { get(st, w, l, false); }
    __device__ __forceinline__ void inv(int st, C2<double>* w, int l) const { get(st, w, l, true); }
};

// forward: lane l, register r hold x[l + 64 r] on entry and X[xw_rev(l) + 64 r]
// on exit (radix-4 decimation in frequency; the register index carries the
// digit being transformed)
template <typename T, class TW>
__device__ __forceinline__ void xw_fwd(C2<T>* x, const TW& t, int l) {
    C2<T> w[3];
    Dft<T, 4, -1>::run(x);
    t.fwd(0, w, l);
    xw_twiddle<T>(x, w);
    xw_swap2<5, 4, T>(x, l);
    Dft<T, 4, -1>::run(x);
    t.fwd(1, w, l);
    xw_twiddle<T>(x, w);
    xw_swap2<3, 2, T>(x, l);
    Dft<T, 4, -1>::run(x);
    t.fwd(2, w, l);
    xw_twiddle<T>(x, w);
    xw_swap2<1, 0, T>(x, l);
    Dft<T, 4, -1>::run(x);
}

// inverse (unnormalised): X[xw_rev(l) + 64 r] on entry, x[l + 64 r] on exit
template <typename T, class TW>
__device__ __forceinline__ void xw_inv(C2<T>* x, const TW& t, int l) {
    C2<T> w[3];
    Dft<T, 4, 1>::run(x);
    t.inv(0, w, l);
    xw_twiddle<T>(x, w);
    xw_swap2<1, 0, T>(x, l);
    Dft<T, 4, 1>::run(x);
    t.inv(1, w, l);
    xw_twiddle<T>(x, w);
    xw_swap2<3, 2, T>(x, l);
    Dft<T, 4, 1>::run(x);
    t.inv(2, w, l);
    xw_twiddle<T>(x, w);
    xw_swap2<5, 4, T>(x, l);
    Dft<T, 4, 1>::run(x);
}

// Kernel E of the merged AB2 history for 256-point lines with the
// wavefront-exchange transforms: k_ns_fourier_xs's loop (one wave per x line,
// persistent one-wave blocks, products and Y loaded into LDS a line ahead,
// 16 bytes per lane per load) with no LDS exchange, so each slot is free for
// the next line's loads once read into registers.  The update runs in the
// forward transform's output order: the lane's frequencies are xw_rev(l) +
// 64 r, where Y is read and written; the emitted inverse transforms return
// to natural order.  Twiddles: FP32 in registers, FP64 from a table in LDS.
template <typename T, bool NT, bool E5, bool REV, int LAUX>
__global__ __launch_bounds__(64) void k_ns_fourier_xw(C2<T>* __restrict__ W, C2<T>* __restrict__ N0,
                                                      const C2<T>* __restrict__ twg, int ny, int nz, int nzp,
                                                      long F, const StepCoef<T>* __restrict__ cfs,
                                                      double* __restrict__ divpart) {
    constexpr int N = 256, E = 4, Q = 64;
    constexpr unsigned VB = sizeof(C2<T>);  // bytes per value
    constexpr int VPL = 16 / VB;            // values one lane loads into LDS at a time
    constexpr int PIECE = 64 * VPL;         // values of one load instruction
    constexpr int PIECES = N / PIECE;
    constexpr int STORES = (E5 ? 8 : 9) * E;  // a line's Y and emitted-field stores
    static_assert(3 * PIECES + STORES < 64, "vmcnt range");
    const StepCoef<T> cf = cfs[0];
    __shared__ C2<T> xs[3][N];   // the products of the line
    __shared__ C2<T> n0l[3][N];  // its Y
    __shared__ C2<T> twt[VB == 16 ? N : 1];
    const int l = threadIdx.x;
    typedef std::conditional_t<VB == 16, XwTwLds, XwTw> TW;
    TW tws;
    if constexpr (VB == 16) {
        for (int t = l; t < N; t += 64) twt[t] = twg[t];
        tws.base = lds_offset(&twt[0]);
        __syncthreads();
    } else {
        tws.load(twg, l);
    }
    const int dr = xw_rev(l);
    const long nlines = (long)ny * nzp;
    auto fetch = [&](C2<T>* slots, const C2<T>* src, long line) {
#pragma unroll
        for (int f = 0; f < 3; f++)
#pragma unroll
            for (int k = 0; k < PIECES; k++)
                __builtin_amdgcn_global_load_lds(
                    (const __attribute__((address_space(1))) void*)(src + f * F + line * N + k * PIECE + VPL * l),
                    (__attribute__((address_space(3))) void*)(slots + f * N + k * PIECE), 16, 0, LAUX);
    };
    // positions p0 + 64 r (r = 0..3) of a landed field, one asm block that waits for them
    typedef T t2 __attribute__((ext_vector_type(2)));
    auto read4 = [&](const C2<T>* slot, int p0, C2<T>* x) {
        const unsigned a = lds_offset(slot + p0);
        t2 v0, v1, v2, v3;
        if constexpr (VB == 16)
            asm volatile(
                "ds_read_b128 %0, %4\n\t"
                "ds_read_b128 %1, %4 offset:%5\n\t"
                "ds_read_b128 %2, %4 offset:%6\n\t"
                "ds_read_b128 %3, %4 offset:%7\n\t"
                "s_waitcnt lgkmcnt(0)"
                : "=&v"(v0), "=&v"(v1), "=&v"(v2), "=&v"(v3)
                : "v"(a), "n"(Q * VB), "n"(2 * Q * VB), "n"(3 * Q * VB)
                : "memory");
        else
            asm volatile(
                "ds_read_b64 %0, %4\n\t"
                "ds_read_b64 %1, %4 offset:%5\n\t"
                "ds_read_b64 %2, %4 offset:%6\n\t"
                "ds_read_b64 %3, %4 offset:%7\n\t"
                "s_waitcnt lgkmcnt(0)"
                : "=&v"(v0), "=&v"(v1), "=&v"(v2), "=&v"(v3)
                : "v"(a), "n"(Q * VB), "n"(2 * Q * VB), "n"(3 * Q * VB)
                : "memory");
        x[0] = mk<T>(v0.x, v0.y);
        x[1] = mk<T>(v1.x, v1.y);
        x[2] = mk<T>(v2.x, v2.y);
        x[3] = mk<T>(v3.x, v3.y);
    };
    double dsum = 0;
    auto at = [&](long t) { return REV ? nlines - 1 - t : t; };
    long line = blockIdx.x;
    if (line < nlines) {
        fetch(&xs[0][0], W, at(line));
        fetch(&n0l[0][0], N0, at(line));
    }
    bool first = true;
    for (; line < nlines; line += gridDim.x) {
        const long nxt = line + gridDim.x;
        const bool more = nxt < nlines;
        const long pl = at(line);
        const int jj = (int)(pl % ny), kk = (int)(pl / ny);
        const T k1 = freq<T>(jj, ny), k2 = freq<T>(kk, nz);
        const bool zline = (jj == ny / 2) || (kk == nz / 2);
        const double dw = (nzp < nz && kk > 0) ? 2.0 : 1.0;
        const long off = pl * N;
        // this line's products landed: issued after them are its Y and, from
        // the second line on, the previous line's stores
        if (first)
            asm volatile("s_waitcnt vmcnt(%0)" ::"n"(3 * PIECES) : "memory");
        else
            asm volatile("s_waitcnt vmcnt(%0)" ::"n"(3 * PIECES + STORES) : "memory");
        C2<T> n1[3][E];
#pragma unroll
        for (int d = 0; d < 3; d++) read4(&xs[d][0], l, n1[d]);
        if (more) fetch(&xs[0][0], W, at(nxt));
#pragma unroll
        for (int d = 0; d < 3; d++) xw_fwd<T>(n1[d], tws, l);
        // this line's Y landed: issued after it are the previous line's stores
        // (not on a block's first line) and the next line's products (not on its last)
        if (first) {
            if (more)
                asm volatile("s_waitcnt vmcnt(%0)" ::"n"(3 * PIECES) : "memory");
            else
                asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
        } else {
            if (more)
                asm volatile("s_waitcnt vmcnt(%0)" ::"n"(3 * PIECES + STORES) : "memory");
            else
                asm volatile("s_waitcnt vmcnt(%0)" ::"n"(STORES) : "memory");
        }
        first = false;
        C2<T> n0s[3][E], vn[3][E];
#pragma unroll
        for (int d = 0; d < 3; d++) read4(&n0l[d][0], dr, n0s[d]);
        if (more) fetch(&n0l[0][0], N0, at(nxt));
#pragma unroll
        for (int s = 0; s < E; s++) {
            const int pos = dr + Q * s;
            const T k0 = freq<T>(pos, N);
            const bool z = zline || pos == N / 2 ||
                           (cf.dealias == 1 && (3 * fabs(k0) > N || 3 * fabs(k1) > ny || 3 * fabs(k2) > nz));
            const T sig = cf.dealias == 2 ? expfilter<T>(k0, N) * expfilter<T>(k1, ny) * expfilter<T>(k2, nz) : T(1);
            C2<T> nn[3], S[3];
#pragma unroll
            for (int d = 0; d < 3; d++) {
                nn[d] = z ? mk<T>(T(0), T(0)) : n1[d][s];
                if (cf.dealias == 2) nn[d] = rmul<T>(sig, nn[d]);
                // S = Y + K/2*(3.*cross1)  (= (V1 + lap1) + F)
                S[d] = cadd<T>(n0s[d][s], rmul<T>(cf.Kh, rmul<T>(T(3), nn[d])));
            }
            const T q = -(k0 * k0 + k1 * k1 + k2 * k2);
            // G = div(S); G.laplacian_inverse() (a=0, b=1)
            C2<T> G = mk<T>(T(0), T(0));
            G = cadd<T>(G, imul<T>(k0, S[0]));
            G = cadd<T>(G, imul<T>(k1, S[1]));
            G = cadd<T>(G, imul<T>(k2, S[2]));
            T sum = 0;
            sum += k0 * k0;
            sum += k1 * k1;
            sum += k2 * k2;
            const T rg = sum != 0 ? T(1) / (T(0) - T(1) * sum) : T(1);
            G.x *= rg;
            G.y *= rg;
            const double rv = 1.0 / (1.0 - (-cf.c1d) * (double)sum);  // 1 where sum == 0
            const T kk3[3] = {k0, k1, k2};
#pragma unroll
            for (int d = 0; d < 3; d++) {
                // V = Vhalf + lap1 = S - G.gradient(), then the CN divide (FP64 factor)
                const C2<T> vv = csub<T>(S[d], imul<T>(kk3[d], G));
                vn[d][s] = mk<T>(T((double)vv.x * rv), T((double)vv.y * rv));
            }
            C2<T> dv = mk<T>(T(0), T(0));
            dv = cadd<T>(dv, imul<T>(k0, vn[0][s]));
            dv = cadd<T>(dv, imul<T>(k1, vn[1][s]));
            dv = cadd<T>(dv, imul<T>(k2, vn[2][s]));
            dsum += dw * sqrt((double)dv.x * (double)dv.x + (double)dv.y * (double)dv.y);
#pragma unroll
            for (int d = 0; d < 3; d++) {
                // the next step's Y = (V + lap V) - K/2 cross1
                const C2<T> yn = csub<T>(cadd<T>(vn[d][s], rmul<T>(cf.c1, rmul<T>(q, vn[d][s]))),
                                         rmul<T>(cf.Kh, nn[d]));
                st_line<NT, T>(N0 + d * F + off + pos, yn);
            }
        }
        // the next step's kernel A: inverse x of V and of i k x V (E5: of
        // i kx Vy and i kx Vz), stored in natural order
#pragma unroll
        for (int g = 0; g < (E5 ? 5 : 6); g++) {
            C2<T> x[E];
#pragma unroll
            for (int s = 0; s < E; s++) {
                const T k0 = freq<T>(dr + Q * s, N);
                if (g < 3)
                    x[s] = vn[g][s];
                else if (E5)
                    x[s] = imul<T>(k0, vn[g - 2][s]);  // d/dx Vy, d/dx Vz
                else if (g == 3)
                    x[s] = csub<T>(imul<T>(k1, vn[2][s]), imul<T>(k2, vn[1][s]));
                else if (g == 4)
                    x[s] = csub<T>(imul<T>(k2, vn[0][s]), imul<T>(k0, vn[2][s]));
                else
                    x[s] = csub<T>(imul<T>(k0, vn[1][s]), imul<T>(k1, vn[0][s]));
            }
            xw_inv<T>(x, tws, l);
#pragma unroll
            for (int s = 0; s < E; s++) st_line<NT, T>(W + g * F + off + l + Q * s, x[s]);
        }
    }
    // this block's sum of |div V| over its lines (FP64; lanes in a fixed order)
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) dsum += __shfl_xor(dsum, o, 64);
    if (l == 0) divpart[blockIdx.x] = dsum;
}

// B of a step whose kernel E emitted five fields (E5): the y inverse of V
// (fields 0..2) and of omega = i k x V, formed on load from V and from the
// x-inverses of i kx Vy and i kx Vz (fields 3, 4), since after the x pass
// omega's other terms need only ky (the line position) and kz (the plane):
//     omega_x = i (ky Vz - kz Vy),  omega_y = i kz Vx - d/dx Vz,
//     omega_z = d/dx Vy - i ky Vx   -> fields 3, 4, 5.
// A block holds its columns of all five inputs in registers and transforms
// the six outputs in turn: 5 field reads and 6 writes instead of 6 and 6.
// In place: each block reads its own tile of every input before writing.
// Its barriers order LDS only (SyncLds): a __syncthreads fence would wait for
// the previous output's stores before every transform.  UNR: the six outputs
// unrolled (more registers, the compiler interleaves them).  RT: the stage
// twiddles held in registers (loaded once) and interleaved-complex exchanges.
// ORD: the inputs loaded in stages, in the order the outputs consume them, so
// that the loads of later fields are in flight while the first outputs are
// transformed; 1: at most three input fields live,
//     load Vx, dVz, dVy -> Vx, omega_y (dVz dead) -> load Vy -> omega_z (Vx,
//     dVy dead) -> load Vz -> Vy, omega_x, Vz;
// 2: at most two,
//     load Vx, dVz -> Vx, omega_y -> load dVy -> omega_z -> load Vy, Vz ->
//     Vy, omega_x, Vz;
// 3: as 2, each load issued once the output that frees its registers is
// formed, before that output's transform (in flight during it).
// OCC: the waves per SIMD the register allocation must allow (two blocks of
// the fewer registers per compute unit, one loading while the other computes).
template <typename T, class P, int LPB, bool NT, bool UNR, bool RT = false, bool LNT = false, int ORD = 0,
          int OCC = 1>
__global__ __launch_bounds__(P::TPL* LPB) __attribute__((amdgpu_waves_per_eu(OCC))) void k_ns_y_curl(
    C2<T>* __restrict__ W, const C2<T>* __restrict__ tw, int ncols, int nouter, long odist, long es, long F, int ny,
    int nz) {
    constexpr int CS = LPB + 1;
    __shared__ __align__(16) T sre[RT ? 2 * P::N * LPB : P::N * CS];
    __shared__ T sim[RT ? 1 : P::N * CS];
    C2<T>* const sx = reinterpret_cast<C2<T>*>(sre);
    const int c = threadIdx.x % LPB;
    const int tid = threadIdx.x / LPB;
    const int tiles = (ncols + LPB - 1) / LPB;
    const int o = blockIdx.x / tiles;
    const int col = (blockIdx.x % tiles) * LPB + c;
    const bool valid = col < ncols && o < nouter;
    C2<T>* base = W + (valid ? (col + (long)o * odist) : 0);
    const T kz = freq<T>(o, nz);
    // 32-bit element offsets within the five (six) fields
    const unsigned uF = (unsigned)F, ues = (unsigned)es;
    C2<T> x[P::E];
    typedef T t2 __attribute__((ext_vector_type(2)));
    // this thread's stage-0 values of input field f
    auto ld = [&](C2<T>* v, int f) {
        for_in<P>(tid, [&](int s, int pos) {
            const C2<T>* a = base + ((unsigned)f * uF + (unsigned)pos * ues);
            if constexpr (LNT) {  // nontemporal loads (read once)
                const t2 w = __builtin_nontemporal_load(reinterpret_cast<const t2*>(a));
                v[s] = mk<T>(w.x, w.y);
            } else {
                v[s] = *a;
            }
        });
    };
    auto idx = [&](int pos) { return RT ? pos * LPB + c : pos * CS + c; };
    C2<T> tws[RT ? rt_size<P>() : 1];
    if constexpr (RT) rt_load<P, 0, T>(tws, tid, tw);
    if constexpr (ORD > 0) {
        // transform x and store it as output field f (a barrier before each
        // transform but the first: the exchange buffer is reused)
        auto emit = [&](int f, bool first) {
            if (!first) SyncLds::run();
            if constexpr (RT)
                line_fft_cr<P, 1, T, decltype(idx), SyncLds>(x, tid, sx, idx, tws);
            else
                line_fft<P, 1, T, decltype(idx), SyncLds>(x, tid, sre, sim, idx, tw);
            if (valid)
                for_out<P>(tid, [&](int s, int pos) {
                    st_line<NT, T>(base + ((unsigned)f * uF + (unsigned)pos * ues), x[s]);
                });
        };
        // ORD 1 -- a: Vx; b: d/dx Vz, then Vy; e: d/dx Vy, then Vz
        // ORD 2 -- a: Vx, then Vy; b: d/dx Vz, then d/dx Vy, then Vz
        C2<T> a[P::E], b[P::E], e[ORD == 1 ? P::E : 1];
        C2<T>* const dvy = ORD == 1 ? e : b;  // d/dx Vy
        C2<T>* const vy = ORD == 1 ? b : a;
        C2<T>* const vz = ORD == 1 ? e : b;
        ld(a, 0);
        ld(b, 4);
        if constexpr (ORD == 1) ld(e, 3);
#pragma unroll
        for (int s = 0; s < P::E; s++) x[s] = a[s];
        emit(0, true);
        constexpr bool EARLY = ORD == 3;
        for_in<P>(tid, [&](int s, int pos) { x[s] = csub<T>(imul<T>(kz, a[s]), b[s]); });  // omega_y
        if constexpr (EARLY) ld(dvy, 3);
        emit(4, false);
        if constexpr (!EARLY) ld(ORD == 1 ? b : dvy, ORD == 1 ? 1 : 3);
        for_in<P>(tid, [&](int s, int pos) { x[s] = csub<T>(dvy[s], imul<T>(freq<T>(pos, ny), a[s])); });  // omega_z
        if constexpr (EARLY) {
            ld(vy, 1);
            ld(vz, 2);
        }
        emit(5, false);
        if constexpr (ORD == 2) ld(vy, 1);
        if constexpr (!EARLY) ld(vz, 2);
#pragma unroll
        for (int s = 0; s < P::E; s++) x[s] = vy[s];
        emit(1, false);
        for_in<P>(tid, [&](int s, int pos) {  // omega_x
            x[s] = csub<T>(imul<T>(freq<T>(pos, ny), vz[s]), imul<T>(kz, vy[s]));
        });
        emit(3, false);
#pragma unroll
        for (int s = 0; s < P::E; s++) x[s] = vz[s];
        emit(2, false);
        return;
    }
    C2<T> v[5][P::E];
#pragma unroll
    for (int f = 0; f < 5; f++) ld(v[f], f);
    // one output at a time; not unrolled, the inputs stay the only live state
#pragma unroll(UNR ? 6 : 1)
    for (int g = 0; g < 6; g++) {
        const int f = g < 3 ? g + 3 : g - 3;  // the vorticity first
        for_in<P>(tid, [&](int s, int pos) {
            const T ky = freq<T>(pos, ny);
            if (f == 3)
                x[s] = csub<T>(imul<T>(ky, v[2][s]), imul<T>(kz, v[1][s]));
            else if (f == 4)
                x[s] = csub<T>(imul<T>(kz, v[0][s]), v[4][s]);
            else if (f == 5)
                x[s] = csub<T>(v[3][s], imul<T>(ky, v[0][s]));
            else
                x[s] = f == 0 ? v[0][s] : f == 1 ? v[1][s] : v[2][s];
        });
        if (g > 0) SyncLds::run();
        if constexpr (RT)
            line_fft_cr<P, 1, T, decltype(idx), SyncLds>(x, tid, sx, idx, tws);
        else
            line_fft<P, 1, T, decltype(idx), SyncLds>(x, tid, sre, sim, idx, tw);
        if (valid)
            for_out<P>(tid, [&](int s, int pos) {
                st_line<NT, T>(base + ((unsigned)f * uF + (unsigned)pos * ues), x[s]);
            });
    }
}

// ---------------------------------------------------------------- conversion kernels
__global__ void k_z_to_c(const double2* __restrict__ a, float2* __restrict__ b, long n) {
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        b[i] = make_float2((float)a[i].x, (float)a[i].y);
}
// The merged AB2 history of kernel E (YS) and back, over the stored planes
// (nzp) of the three fields of box blockIdx.y (field stride F, box offset
// boff), with that box's coefficients: TOY: H = (V + nu K/2 lap V) - K/2 H
// (H = N0 in, Y out); otherwise H = ((V + nu K/2 lap V) - H) / (K/2) (Y in,
// N0 out)
template <typename T, bool TOY>
__global__ void k_yform(const C2<T>* __restrict__ V, C2<T>* __restrict__ H, long F, long boff, int nx, int ny,
                        int nz, int nzp, const StepCoef<T>* __restrict__ cfs) {
    const long n = (long)nx * ny * nzp;
    const T c1 = cfs[blockIdx.y].c1, Kh = cfs[blockIdx.y].Kh;
    V += blockIdx.y * boff;
    H += blockIdx.y * boff;
    for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < 3 * n; e += (long)gridDim.x * blockDim.x) {
        const int d = (int)(e / n);
        const long r = e % n;
        const long jk = r / nx;
        const T k0 = freq<T>((int)(r % nx), nx), k1 = freq<T>((int)(jk % ny), ny), k2 = freq<T>((int)(jk / ny), nz);
        const T q = -(k0 * k0 + k1 * k1 + k2 * k2);
        const long o = d * F + r;
        const C2<T> v = V[o], h = H[o];
        const C2<T> a = cadd<T>(v, rmul<T>(c1, rmul<T>(q, v)));  // V1 + lap1
        H[o] = TOY ? csub<T>(a, rmul<T>(Kh, h)) : mk<T>((a.x - h.x) / Kh, (a.y - h.y) / Kh);
    }
}

__global__ void k_c_to_z(const float2* __restrict__ a, double2* __restrict__ b, long n) {
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        b[i] = make_double2((double)a[i].x, (double)a[i].y);
}
__global__ void k_abs_sum_partial(const double* __restrict__ part, int np, double* __restrict__ out) {
    __shared__ double s[256];
    double acc = 0;
    for (int i = threadIdx.x; i < np; i += 256) acc += part[i];
    s[threadIdx.x] = acc;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = s[0];
}

// Box filter of ratio r on physical velocity u (3 fields, real parts): each
// coarse cell gets ubar_i = <u_i> and the subgrid stress
// tau_ij = <u_i u_j> - ubar_i ubar_j over its r^3 fine cells.  Output: 9
// float fields (ubar x y z, tau xx yy zz xy xz yz) of nc coarse cells.
// Box b is blockIdx.y: its field d is at u + (d*nboxes + b)*N (field stride
// fs = nboxes*N) and its output at out + b*9*nc.
template <typename T>
__global__ void k_box_filter(const C2<T>* __restrict__ u, long N, long fs, int nx, int ny, int r,
                             int cx, int cy, long nc, float* __restrict__ out) {
    u += blockIdx.y * N;
    out += blockIdx.y * 9 * nc;
    for (long c = blockIdx.x * (long)blockDim.x + threadIdx.x; c < nc; c += (long)gridDim.x * blockDim.x) {
        const int I = (int)(c % cx);
        const long rr = c / cx;
        const int J = (int)(rr % cy);
        const int K = (int)(rr / cy);
        double s[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        for (int dk = 0; dk < r; dk++)
            for (int dj = 0; dj < r; dj++)
                for (int di = 0; di < r; di++) {
                    const long n = (long)(I * r + di) + (long)nx * ((long)(J * r + dj) + (long)ny * (K * r + dk));
                    const double a = (double)u[n].x, b = (double)u[fs + n].x, d = (double)u[2 * fs + n].x;
                    s[0] += a; s[1] += b; s[2] += d;
                    s[3] += a * a; s[4] += b * b; s[5] += d * d;
                    s[6] += a * b; s[7] += a * d; s[8] += b * d;
                }
        const double inv = 1.0 / ((double)r * r * r);
        const double ua = s[0] * inv, ub = s[1] * inv, ud = s[2] * inv;
        out[c] = (float)ua;
        out[nc + c] = (float)ub;
        out[2 * nc + c] = (float)ud;
        out[3 * nc + c] = (float)(s[3] * inv - ua * ua);
        out[4 * nc + c] = (float)(s[4] * inv - ub * ub);
        out[5 * nc + c] = (float)(s[5] * inv - ud * ud);
        out[6 * nc + c] = (float)(s[6] * inv - ua * ub);
        out[7 * nc + c] = (float)(s[7] * inv - ua * ud);
        out[8 * nc + c] = (float)(s[8] * inv - ub * ud);
    }
}

// ratio-r box averages of the physical u x w (3 fields of each box; u and w
// field-major with field stride fs, box = blockIdx.y): out[box][3][nc]
template <typename T>
__global__ void k_box_cross(const C2<T>* __restrict__ u, const C2<T>* __restrict__ w, long N, long fs, int nx,
                            int ny, int r, int cx, int cy, long nc, float* __restrict__ out) {
    u += blockIdx.y * N;
    w += blockIdx.y * N;
    out += blockIdx.y * 3 * nc;
    for (long c = blockIdx.x * (long)blockDim.x + threadIdx.x; c < nc; c += (long)gridDim.x * blockDim.x) {
        const int I = (int)(c % cx);
        const long rr = c / cx;
        const int J = (int)(rr % cy);
        const int K = (int)(rr / cy);
        double s[3] = {0, 0, 0};
        for (int dk = 0; dk < r; dk++)
            for (int dj = 0; dj < r; dj++)
                for (int di = 0; di < r; di++) {
                    const long n = (long)(I * r + di) + (long)nx * ((long)(J * r + dj) + (long)ny * (K * r + dk));
                    const double u0 = u[n].x, u1 = u[fs + n].x, u2 = u[2 * fs + n].x;
                    const double w0 = w[n].x, w1 = w[fs + n].x, w2 = w[2 * fs + n].x;
                    s[0] += u1 * w2 - u2 * w1;
                    s[1] += u2 * w0 - u0 * w2;
                    s[2] += u0 * w1 - u1 * w0;
                }
        const double inv = 1.0 / ((double)r * r * r);
        for (int a = 0; a < 3; a++) out[a * nc + c] = (float)(s[a] * inv);
    }
}

// w = i k x v for the field-major layout (box = blockIdx.y)
template <typename T>
__global__ void k_curl_boxes(const C2<T>* __restrict__ v, C2<T>* __restrict__ w, long N, long fs, int nx, int ny,
                             int nz) {
    v += blockIdx.y * N;
    w += blockIdx.y * N;
    for (long p = blockIdx.x * (long)blockDim.x + threadIdx.x; p < N; p += (long)gridDim.x * blockDim.x) {
        const int i = (int)(p % nx);
        const long rr = p / nx;
        const T k0 = freq<T>(i, nx), k1 = freq<T>((int)(rr % ny), ny), k2 = freq<T>((int)(rr / ny), nz);
        const C2<T> a = v[p], b = v[fs + p], c = v[2 * fs + p];
        w[p] = csub<T>(imul<T>(k1, c), imul<T>(k2, b));
        w[fs + p] = csub<T>(imul<T>(k2, a), imul<T>(k0, c));
        w[2 * fs + p] = csub<T>(imul<T>(k0, b), imul<T>(k1, a));
    }
}

static unsigned egrid(long n) {
    long g = (n + 255) / 256;
    if (g > 262144) g = 262144;
    return (unsigned)(g < 1 ? 1 : g);
}

}  // namespace sh

using namespace sh;

extern "C" {
int sh_gradient_z(const void*, void*, void*, void*, int, int, int, void*);
int sh_laplacian_inverse_z(void*, double, double, int, int, int, void*);
int sh_divergence_z(const void*, const void*, const void*, void*, int, int, int, void*);
int sh_curl_z(const void*, const void*, const void*, void*, void*, void*, int, int, int, void*);
int sh_vlaplacian_z(const void*, const void*, const void*, void*, void*, void*, int, int, int, void*);
int sh_cross_z(const void*, const void*, void*, void*, int, int, int, void*);
int sh_add_z(void*, const void*, const void*, long, void*);
int sh_sub_z(void*, const void*, const void*, long, void*);
int sh_scale_z(void*, double, const void*, long, void*);
int sh_average_z(const void*, long, double*, void*);
}

struct sh_ns_s {
    int nx, ny, nz;
    long N;  // elements per field
    double nu, K;
    int prec;    // SH_F64 / SH_F32
    int fused;
    int nonlinear;
    hipStream_t st;
    int device;
    // FP64 op-by-op state (also used for the start-up in every mode)
    double2 *V0, *V1, *om0, *om1, *cr0, *cr1;  // 3 fields each
    double2 *Vn, *omn, *Fv, *lap1, *S, *Gg, *Vh, *work;  // work: 6 fields
    double2 *G, *div, *Pi0;
    // fused state (precision prec): V1f, N0f (3 fields each), Wf (6 fields),
    // for nboxes equal boxes laid out field-major: field f of box b at (f*nboxes + b)*N
    void *V1f, *N0f, *Wf;
    int nboxes;
    std::vector<double> nus;  // viscosity of each box
    void* coefs;              // StepCoef<T>[nboxes] on the device
    void* snap;               // 3 physical fields for snapshots (allocated on first use)
    double* divpart;          // [nboxes][divblocks + 1]
    int divblocks;
    int div_nb;  // partials written by the last kernel E
    bool have_div;
    bool w_primed;  // W holds the x-inverse of V1 and curl V1 for the next step
    // the merged AB2 history (kernel E with YS): N0f holds Y = (V1 + lap1) -
    // K/2 N0, V1f is stale, and the first three W fields hold the unnormalised
    // x-inverse of V1 (W is primed); ns_canonical returns to (V1, N0)
    bool yform;
    bool keep_v1;  // the merged history is off (a caller reads V1 in place)
    // W holds the five-field emission (V and i kx Vy, i kx Vz x-inverted): the
    // next y inverse forms the vorticity on load (k_ns_y_curl)
    bool w5;
    double div_avg;
    const void* forcing;  // body force (3 fields per box, precision prec) or null
    int cur_box;          // box whose start-up is running
    int dealias;          // dealiasing of the nonlinear term (0: none, as the reference)
    int scheme;           // 0: exec/ns.cpp (CN viscous), 1: exec/ns2.cpp (backward Euler viscous)
    int layout;           // 1: half-spectrum layout when the state is Hermitian; 0: always full
    int overlap;          // 1: kernels B-D run as two column halves on two streams
    hipStream_t st2;      // second stream of the overlapped schedule (default priority)
    hipEvent_t ev_fork, ev_join;
    hipStream_t cst;      // stream and column range of the B-D launch being issued
    int cx0, cnx;
    bool half;            // the fused state holds only the planes kz < nz/2
    std::vector<char> box_herm;  // per box: its last start-up state was Hermitian
    std::vector<void*> owned;
};

namespace {

int alloc_fields(sh_ns_s* s, void** p, int nfields, size_t elem) {
    int rc = pool_alloc(p, elem * (size_t)s->N * nfields);
    if (rc) return rc;
    s->owned.push_back(*p);
    return hipMemsetAsync(*p, 0, elem * (size_t)s->N * nfields, s->st) == hipSuccess ? SH_OK : SH_EHIP;
}

#define F3(p, d) ((p) + (long)(d) * s->N)

int vlap(sh_ns_s* s, const double2* v, double2* o) {
    return sh_vlaplacian_z(F3(v, 0), F3(v, 1), F3(v, 2), F3(o, 0), F3(o, 1), F3(o, 2), s->nx, s->ny,
                           s->nz, s->st);
}
int curl(sh_ns_s* s, const double2* v, double2* o) {
    return sh_curl_z(F3(v, 0), F3(v, 1), F3(v, 2), F3(o, 0), F3(o, 1), F3(o, 2), s->nx, s->ny, s->nz,
                     s->st);
}
int cross(sh_ns_s* s, const double2* a, const double2* b, double2* o) {
    if (!s->nonlinear) {
        return hipMemsetAsync(o, 0, sizeof(double2) * 3 * s->N, s->st) == hipSuccess ? SH_OK : SH_EHIP;
    }
    return sh_cross_z(a, b, o, s->work, s->nx, s->ny, s->nz, s->st);
}

#define RC(x)                \
    do {                     \
        int _rc = (x);       \
        if (_rc) return _rc; \
    } while (0)

// dealiasing of 3 fields as the fused kernel E does it: mode 1 the 2/3-rule
// truncation, mode 2 the exponential filter
__global__ void k_dealias3(double2* __restrict__ f, long fs, int nx, int ny, int nz, int mode) {
    const long N = (long)nx * ny * nz;
    for (long p = blockIdx.x * (long)blockDim.x + threadIdx.x; p < N; p += (long)gridDim.x * blockDim.x) {
        const int i = (int)(p % nx), j = (int)((p / nx) % ny), k = (int)(p / ((long)nx * ny));
        const double k0 = freq<double>(i, nx), k1 = freq<double>(j, ny), k2 = freq<double>(k, nz);
        if (mode == 1) {
            if (3 * fabs(k0) > nx || 3 * fabs(k1) > ny || 3 * fabs(k2) > nz)
                for (int d = 0; d < 3; d++) f[d * fs + p] = make_double2(0.0, 0.0);
        } else {
            const double s = expfilter<double>(k0, nx) * expfilter<double>(k1, ny) * expfilter<double>(k2, nz);
            for (int d = 0; d < 3; d++) f[d * fs + p] = make_double2(s * f[d * fs + p].x, s * f[d * fs + p].y);
        }
    }
}

int dealias64(sh_ns_s* s, double2* cr) {
    if (!s->dealias) return SH_OK;
    hipLaunchKernelGGL(k_dealias3, dim3(egrid(s->N)), dim3(256), 0, s->st, cr, s->N, s->nx, s->ny, s->nz,
                       s->dealias);
    SH_LAUNCH_CHECK("k_dealias3");
    return SH_OK;
}

// cr += the body force of the current box (an FP32 force is widened through work)
int add_forcing64(sh_ns_s* s, double2* cr) {
    if (!s->forcing) return SH_OK;
    const long N = s->N;
    for (int d = 0; d < 3; d++) {
        const long off = ((long)d * s->nboxes + s->cur_box) * N;
        const void* f = (const double2*)s->forcing + off;
        if (s->prec == SH_F32) {
            hipLaunchKernelGGL(k_c_to_z, dim3(egrid(N)), dim3(256), 0, s->st, (const float2*)s->forcing + off,
                               F3(s->work, d), N);
            SH_LAUNCH_CHECK("k_c_to_z");
            f = F3(s->work, d);
        }
        RC(sh_add_z(F3(cr, d), F3(cr, d), f, N, s->st));
    }
    return SH_OK;
}

// exec/ns.cpp:132-153
// ---------------------------------------------------------------- merged AB2 history
// The merged history serves one FP64 box with the Crank-Nicolson scheme and no
// body force (kernel E with YS); everything else keeps (V1, N0).
bool y_ok(const sh_ns_s* s) {
    return s->fused && !s->forcing && s->scheme == 0 && !s->keep_v1;
}

bool capturing(hipStream_t st) {
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    return hipStreamIsCapturing(st, &cs) == hipSuccess && cs != hipStreamCaptureStatusNone;
}

// V1 of a solver in the merged form into dst (3 fields of every box, at
// stride N in the state's order: field d of box b in slot d * nboxes + b):
// the forward x pass of W's first three fields (their unnormalised
// x-inverse), scaled by 1/nx, Nyquist lines zeroed as every forward() leaves them
template <typename T>
int v1_into(sh_ns_s* s, C2<T>* dst) {
    const int nzp = s->half ? s->nz / 2 : s->nz;
    const size_t cnt = (size_t)s->nx * s->ny * nzp;
    const int slots = 3 * s->nboxes;
    for (int q = 0; q < slots; q++)
        SH_HIP(hipMemcpyAsync(dst + q * s->N, (const C2<T>*)s->Wf + q * s->N, sizeof(C2<T>) * cnt,
                              hipMemcpyDeviceToDevice, s->st));
    Epi<T> ep{T(1.0 / s->nx), 3, s->ny, s->ny / 2, s->nz / 2};
    return pass_contig<T>(dst, s->nx, (long)s->ny * nzp, s->nx, slots, s->N, -1, ep, s->st, -1);
}

// the same for box b only, into dst (3 fields at stride N)
template <typename T>
int v1_into_box(sh_ns_s* s, int b, C2<T>* dst) {
    const int nzp = s->half ? s->nz / 2 : s->nz;
    const size_t cnt = (size_t)s->nx * s->ny * nzp;
    for (int d = 0; d < 3; d++)
        SH_HIP(hipMemcpyAsync(dst + d * s->N, (const C2<T>*)s->Wf + ((long)d * s->nboxes + b) * s->N,
                              sizeof(C2<T>) * cnt, hipMemcpyDeviceToDevice, s->st));
    Epi<T> ep{T(1.0 / s->nx), 3, s->ny, s->ny / 2, s->nz / 2};
    return pass_contig<T>(dst, s->nx, (long)s->ny * nzp, s->nx, 3, s->N, -1, ep, s->st, -1);
}

// TOY: N0 -> Y; otherwise Y -> N0 (V1 current), with the step's coefficients
template <typename T, bool TOY>
int launch_yform(sh_ns_s* s) {
    const int nzp = s->half ? s->nz / 2 : s->nz;
    hipLaunchKernelGGL((k_yform<T, TOY>), dim3(egrid(3L * s->nx * s->ny * nzp), (unsigned)s->nboxes), dim3(256), 0,
                       s->st, (const C2<T>*)s->V1f, (C2<T>*)s->N0f, s->N * s->nboxes, s->N, s->nx, s->ny, s->nz,
                       nzp, (const StepCoef<T>*)s->coefs);
    SH_LAUNCH_CHECK("k_yform");
    return SH_OK;
}

// back to (V1, N0): V1 from W, N0 = ((V1 + lap1) - Y) / (K/2).  Never inside
// a graph capture: a replay would convert again a state already converted.
int ns_canonical(sh_ns_s* s) {
    if (!s->yform) return SH_OK;
    if (capturing(s->st)) {
        set_error("ns: the solver state would change form inside a graph capture; run one step "
                  "with the captured schedule before capturing");
        return SH_EINVAL;
    }
    if (s->prec == SH_F64) {
        RC(v1_into<double>(s, (double2*)s->V1f));
        RC((launch_yform<double, false>(s)));
    } else {
        RC(v1_into<float>(s, (float2*)s->V1f));
        RC((launch_yform<float, false>(s)));
    }
    s->yform = false;
    return SH_OK;
}

int ensure_snap(sh_ns_s* s);

// V1 of box b of a merged-history state as complex128 into out (3 fields at
// stride N), leaving the state as it is (FP32: through the snapshot scratch)
int v1_into_z(sh_ns_s* s, double2* out, int b = 0) {
    if (s->prec == SH_F64) return v1_into_box<double>(s, b, out);
    RC(ensure_snap(s));
    RC(v1_into_box<float>(s, b, (float2*)s->snap));
    const long n3 = 3 * s->N;
    hipLaunchKernelGGL(k_c_to_z, dim3(egrid(n3)), dim3(256), 0, s->st, (const float2*)s->snap, out, n3);
    SH_LAUNCH_CHECK("k_c_to_z");
    return SH_OK;
}

int startup64(sh_ns_s* s) {
    const long V = 3 * s->N;
    RC(cross(s, s->V0, s->om0, s->cr1));
    RC(add_forcing64(s, s->cr1));
    RC(dealias64(s, s->cr1));
    RC(vlap(s, s->V0, s->lap1));
    RC(sh_scale_z(s->lap1, s->nu, s->lap1, V, s->st));
    RC(sh_add_z(s->S, s->cr1, s->lap1, V, s->st));
    RC(sh_divergence_z(F3(s->S, 0), F3(s->S, 1), F3(s->S, 2), s->Pi0, s->nx, s->ny, s->nz, s->st));
    RC(sh_laplacian_inverse_z(s->Pi0, 0.0, 1.0, s->nx, s->ny, s->nz, s->st));
    RC(sh_gradient_z(s->Pi0, F3(s->Gg, 0), F3(s->Gg, 1), F3(s->Gg, 2), s->nx, s->ny, s->nz, s->st));
    RC(sh_sub_z(s->S, s->S, s->Gg, V, s->st));
    RC(sh_scale_z(s->S, s->K, s->S, V, s->st));
    RC(sh_add_z(s->V1, s->V0, s->S, V, s->st));
    RC(curl(s, s->V1, s->om1));
    RC(cross(s, s->V0, s->om0, s->cr0));
    RC(add_forcing64(s, s->cr0));
    RC(dealias64(s, s->cr0));
    return SH_OK;
}

// exec/ns.cpp:167-217
int step64(sh_ns_s* s) {
    const long V = 3 * s->N;
    const double K = s->K, nu = s->nu;
    const int nx = s->nx, ny = s->ny, nz = s->nz;
    RC(cross(s, s->V1, s->om1, s->cr1));
    RC(add_forcing64(s, s->cr1));
    RC(dealias64(s, s->cr1));
    RC(sh_scale_z(s->Fv, 3.0, s->cr1, V, s->st));
    RC(sh_sub_z(s->Fv, s->Fv, s->cr0, V, s->st));
    RC(sh_scale_z(s->Fv, K / 2, s->Fv, V, s->st));
    // exec/ns2.cpp:178, 185-189: lap1 = nu*K*V1.laplacian(); V = Vhalf; divide by 1 + nu K |k|^2
    const double c1 = s->scheme ? nu * K : nu * K / 2;
    RC(vlap(s, s->V1, s->lap1));
    RC(sh_scale_z(s->lap1, c1, s->lap1, V, s->st));
    RC(sh_add_z(s->S, s->V1, s->lap1, V, s->st));
    RC(sh_add_z(s->S, s->S, s->Fv, V, s->st));
    RC(sh_divergence_z(F3(s->S, 0), F3(s->S, 1), F3(s->S, 2), s->G, nx, ny, nz, s->st));
    RC(sh_laplacian_inverse_z(s->G, 0.0, 1.0, nx, ny, nz, s->st));
    RC(sh_gradient_z(s->G, F3(s->Gg, 0), F3(s->Gg, 1), F3(s->Gg, 2), nx, ny, nz, s->st));
    RC(sh_add_z(s->Vh, s->V1, s->Fv, V, s->st));
    RC(sh_sub_z(s->Vh, s->Vh, s->Gg, V, s->st));
    if (s->scheme)
        SH_HIP(hipMemcpyAsync(s->Vn, s->Vh, sizeof(double2) * V, hipMemcpyDeviceToDevice, s->st));
    else
        RC(sh_add_z(s->Vn, s->Vh, s->lap1, V, s->st));
    for (int d = 0; d < 3; d++) RC(sh_laplacian_inverse_z(F3(s->Vn, d), 1.0, -c1, nx, ny, nz, s->st));
    RC(curl(s, s->Vn, s->omn));
    RC(sh_divergence_z(F3(s->Vn, 0), F3(s->Vn, 1), F3(s->Vn, 2), s->div, nx, ny, nz, s->st));
    s->have_div = true;
    // histories: V0 = V1; V1 = V; omega0 = omega1; omega1 = omega; cross0 = cross1
    SH_HIP(hipMemcpyAsync(s->V0, s->V1, sizeof(double2) * V, hipMemcpyDeviceToDevice, s->st));
    SH_HIP(hipMemcpyAsync(s->V1, s->Vn, sizeof(double2) * V, hipMemcpyDeviceToDevice, s->st));
    SH_HIP(hipMemcpyAsync(s->om0, s->om1, sizeof(double2) * V, hipMemcpyDeviceToDevice, s->st));
    SH_HIP(hipMemcpyAsync(s->om1, s->omn, sizeof(double2) * V, hipMemcpyDeviceToDevice, s->st));
    SH_HIP(hipMemcpyAsync(s->cr0, s->cr1, sizeof(double2) * V, hipMemcpyDeviceToDevice, s->st));
    return SH_OK;
}

// ---------------------------------------------------------------- fused launchers
template <typename T, class P>
int launch_inv_x(sh_ns_s* s) {
    constexpr int LPB = (512 / (6 * P::TPL)) >= 8 ? 8 : (512 / (6 * P::TPL)) >= 4 ? 4 : (512 / (6 * P::TPL)) >= 2 ? 2 : 1;
    const C2<T>* tw = (const C2<T>*)twiddles(P::N, sizeof(T) == 8 ? SH_F64 : SH_F32);
    if (!tw) return SH_ENOMEM;
    const int nzp = s->half ? s->nz / 2 : s->nz;
    const long nlines = (long)s->ny * nzp;
    dim3 grid((unsigned)((nlines + LPB - 1) / LPB), (unsigned)s->nboxes), block(P::TPL * LPB * 6);
    ProfScope ps("ns_inv_x", s->st);
    hipLaunchKernelGGL((k_ns_inv_x<T, P, LPB>), grid, block, 0, s->st, (const C2<T>*)s->V1f,
                       (C2<T>*)s->Wf, tw, s->ny, s->nz, nzp, s->N * s->nboxes, s->N);
    SH_LAUNCH_CHECK("k_ns_inv_x");
    s->w5 = false;
    return SH_OK;
}

// the y inverse that forms the vorticity on load (k_ns_y_curl) over the
// columns [cx0, cx0 + cnx) on stream cst
template <typename T, class P, int LPB, bool NT = true, bool UNR = false, bool RT = false, bool LNT = false,
          int ORD = 0, int OCC = 1>
int launch_y_curl(sh_ns_s* s) {
    const C2<T>* tw = (const C2<T>*)twiddles(P::N, sizeof(T) == 8 ? SH_F64 : SH_F32);
    if (!tw) return SH_ENOMEM;
    if (P::N != s->ny || s->nboxes != 1) {
        set_error("ns: the vorticity-forming y pass needs one box and a plan of length ny");
        return SH_EINVAL;
    }
    const int nzp = s->half ? s->nz / 2 : s->nz;
    const long tiles = (s->cnx + LPB - 1) / LPB;
    ProfScope ps("ns_y_inv", s->cst);
    hipLaunchKernelGGL((k_ns_y_curl<T, P, LPB, NT, UNR, RT, LNT, ORD, OCC>), dim3((unsigned)(tiles * nzp)),
                       dim3(P::TPL * LPB), 0,
                       s->cst,
                       (C2<T>*)s->Wf + s->cx0, tw, s->cnx, nzp, (long)s->nx * s->ny, (long)s->nx, s->N, s->ny,
                       s->nz);
    SH_LAUNCH_CHECK("k_ns_y_curl");
    return SH_OK;
}

template <typename T, class P, int LPB>
int launch_cross_z(sh_ns_s* s) {
    const C2<T>* tw = (const C2<T>*)twiddles(P::N, sizeof(T) == 8 ? SH_F64 : SH_F32);
    if (!tw) return SH_ENOMEM;
    if (6 * s->N * s->nboxes > 0xffffffffL) {  // the kernel addresses W with 32-bit element offsets
        set_error("ns: grid too large for the fused z kernel");
        return SH_EINVAL;
    }
    const int tiles = (s->cnx + LPB - 1) / LPB;
    dim3 grid((unsigned)(tiles * s->ny), (unsigned)s->nboxes), block(P::TPL * LPB);
    const T scale = T(1.0 / ((double)s->nx * (double)s->ny * (double)s->nz));
    ProfScope ps("ns_cross_z", s->cst);
    hipLaunchKernelGGL((k_ns_cross_z<T, P, LPB>), grid, block, 0, s->cst, (C2<T>*)s->Wf + s->cx0, tw, s->nx,
                       s->cnx, s->ny, s->N * s->nboxes, scale, s->N);
    SH_LAUNCH_CHECK("k_ns_cross_z");
    return SH_OK;
}

// kernel C of the half-spectrum layout: P is the length-nz/2 plan
template <typename T, class P, int LPB, bool NT = false>
int launch_cross_zh(sh_ns_s* s) {
    const int prec = sizeof(T) == 8 ? SH_F64 : SH_F32;
    const C2<T>* tw = (const C2<T>*)twiddles(P::N, prec);
    const C2<T>* twn = (const C2<T>*)twiddles(s->nz, prec);
    if (!tw || !twn) {
        set_error("ns: twiddle table unavailable (first use inside a graph capture?)");
        return SH_ENOMEM;
    }
    if (2 * P::N != s->nz) {
        set_error("ns: half-spectrum z kernel of the wrong length");
        return SH_EINVAL;
    }
    if (6 * s->N * s->nboxes > 0xffffffffL) {  // 32-bit element offsets, as kernel C
        set_error("ns: grid too large for the fused z kernel");
        return SH_EINVAL;
    }
    const int tiles = (s->cnx + LPB - 1) / LPB;
    dim3 grid((unsigned)(tiles * s->ny), (unsigned)s->nboxes), block(P::TPL * LPB);
    const T scale = T(1.0 / ((double)s->nx * (double)s->ny * (double)s->nz));
    ProfScope ps("ns_cross_z", s->cst);
    hipLaunchKernelGGL((k_ns_cross_zh<T, P, LPB, NT>), grid, block, 0, s->cst, (C2<T>*)s->Wf + s->cx0, tw, twn,
                       s->nx, s->cnx, s->ny, s->N * s->nboxes, scale, s->N);
    SH_LAUNCH_CHECK("k_ns_cross_zh");
    return SH_OK;
}

// kernel C of the half-spectrum layout with LDS-streamed inputs (FP64): a
// persistent grid of the resident blocks looping over the column tiles
template <typename T, class P, int LPB, bool NT = false, bool IL = false, bool HALF = true, int RT = 0,
          int LAUX = 0>
int launch_cross_zd(sh_ns_s* s) {
    constexpr int prec = sizeof(T) == 8 ? SH_F64 : SH_F32;
    // an FP32 lane loads two neighbouring columns: an odd column count would
    // read one value past the last column of the grid
    if (sizeof(T) == 4 && (s->cnx % 2) != 0) return launch_cross_zh<T, P, LPB>(s);
    const C2<T>* tw = (const C2<T>*)twiddles(P::N, prec);
    const C2<T>* twn = HALF ? (const C2<T>*)twiddles(s->nz, prec) : tw;
    if (!tw || !twn) {
        set_error("ns: twiddle table unavailable (first use inside a graph capture?)");
        return SH_ENOMEM;
    }
    if ((HALF ? 2 : 1) * P::N != s->nz) {
        set_error("ns: fused z kernel of the wrong length");
        return SH_EINVAL;
    }
    if (6 * s->N * s->nboxes > 0xffffffffL) {  // 32-bit element offsets, as kernel C
        set_error("ns: grid too large for the fused z kernel");
        return SH_EINVAL;
    }
    // resident blocks: CUs x blocks per CU (LDS-bound: one 16-column block,
    // two 8-column ones)
    static int resident = 0;
    if (resident == 0) {
        int dev = 0, v = 0, per = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || v <= 0)
            v = 256;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k_ns_cross_zd<P, LPB, NT, IL, HALF, T, RT, LAUX>,
                                                         P::TPL * LPB,
                                                         0) != hipSuccess ||
            per <= 0)
            per = 1;
        resident = v * per;
    }
    const int tiles = (s->cnx + LPB - 1) / LPB;
    const int ntile = tiles * s->ny;
    const int g = ntile < resident ? ntile : resident;
    dim3 grid((unsigned)g, (unsigned)s->nboxes), block(P::TPL * LPB);
    const T scale = T(1.0 / ((double)s->nx * (double)s->ny * (double)s->nz));
    ProfScope ps("ns_cross_z", s->cst);
    hipLaunchKernelGGL((k_ns_cross_zd<P, LPB, NT, IL, HALF, T, RT, LAUX>), grid, block, 0, s->cst, (C2<T>*)s->Wf + s->cx0,
                       tw, twn, s->nx, s->cnx, s->ny, s->N * s->nboxes, scale, s->N, ntile);
    SH_LAUNCH_CHECK("k_ns_cross_zd");
    return SH_OK;
}

// EMIT = false leaves kernel A to the next step (w_primed false), which moves
// that work off this kernel into the next step's first phase
template <typename T, class P, int LPB, bool EMIT = true, bool BL = false, bool NT = false, bool YS = false,
          bool E5 = false, bool LNT = false>
int launch_fourier_x(sh_ns_s* s) {
    if (!YS && s->yform) {
        set_error("ns: kernel E of the (V1, N0) form on a merged-history state");
        return SH_EINVAL;
    }
    if (s->w5 != E5 && capturing(s->st)) {
        set_error("ns: the emitted form of kernel E would change inside a graph capture; run one step "
                  "with the captured schedule before capturing");
        return SH_EINVAL;
    }
    const C2<T>* tw = (const C2<T>*)twiddles(P::N, sizeof(T) == 8 ? SH_F64 : SH_F32);
    if (!tw) return SH_ENOMEM;
    const int nzp = s->half ? s->nz / 2 : s->nz;
    const long nlines = (long)s->ny * nzp;
    const unsigned nb = (unsigned)((nlines + LPB - 1) / LPB);
    if ((int)nb > s->divblocks) {
        set_error("ns: divergence partial buffer too small");
        return SH_EINVAL;
    }
    ProfScope ps("ns_fourier_x", s->st);
    hipLaunchKernelGGL((k_ns_fourier_x<T, P, LPB, EMIT, BL, NT, YS, E5, LNT>), dim3(nb, (unsigned)s->nboxes),
                       dim3(P::TPL * LPB), 0, s->st, (C2<T>*)s->Wf, (C2<T>*)s->V1f,
                       (C2<T>*)s->N0f, (const C2<T>*)s->forcing, tw, s->ny, s->nz, nzp, s->N * s->nboxes,
                       (const StepCoef<T>*)s->coefs, s->divpart, s->N, s->divblocks + 1);
    SH_LAUNCH_CHECK("k_ns_fourier_x");
    s->div_nb = (int)nb;
    s->w_primed = EMIT;
    s->w5 = E5;
    if (YS) s->yform = true;
    return SH_OK;
}

// the register-loaded kernel E with the merged AB2 history: converts a
// (V1, N0) state first (never inside a graph capture); where the merged form
// does not apply, the (V1, N0) kernel of the same schedule
// a vorticity-forming y pass is compiled for this solver's ny (a five-field
// kernel E needs one for the next step)
bool y_curl_ok(sh_ns_s* s);

template <typename T, class P, int LPB, bool BL = false, bool NT = false, bool E5 = false, bool LNT = false>
int launch_fourier_xm(sh_ns_s* s) {
    const bool cap = capturing(s->st);
    if (!y_ok(s) || (!s->yform && cap)) return launch_fourier_x<T, P, LPB, true, BL, NT, false, false, LNT>(s);
    if (!s->yform) RC((launch_yform<T, true>(s)));
    // inside a capture the emitted form stays what the captured y pass expects
    if (cap ? s->w5 : (E5 && y_curl_ok(s)))
        return launch_fourier_x<T, P, LPB, true, BL, NT, true, true, LNT>(s);
    return launch_fourier_x<T, P, LPB, true, BL, NT, true, false, LNT>(s);
}

// kernel E with LDS-streamed inputs: a persistent grid of one-wave blocks, as
// many as are resident, at most one per line; with a body force the
// register-loaded kernel of the same plan runs instead
template <typename T, class P, bool NT = true, bool SHARE = false, bool G3 = false>
int launch_fourier_xd(sh_ns_s* s) {
    static_assert(sizeof(T) == 8, "FP64 only");
    if (s->forcing) return launch_fourier_x<T, P, 1, true, false, NT>(s);
    if (s->yform) {
        set_error("ns: kernel E of the (V1, N0) form on a merged-history state");
        return SH_EINVAL;
    }
    if (s->w5 && capturing(s->st)) {
        set_error("ns: the emitted form of kernel E would change inside a graph capture; run one step "
                  "with the captured schedule before capturing");
        return SH_EINVAL;
    }
    const C2<T>* tw = (const C2<T>*)twiddles(P::N, SH_F64);
    if (!tw) return SH_ENOMEM;
    const int nzp = s->half ? s->nz / 2 : s->nz;
    const long nlines = (long)s->ny * nzp;
    static int resident = 0;
    if (resident == 0) {
        int dev = 0, v = 0, per = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || v <= 0)
            v = 256;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k_ns_fourier_xd<P, NT, SHARE, G3>, 64, 0) != hipSuccess ||
            per <= 0)
            per = 1;
        resident = v * per;
    }
    const unsigned nb = (unsigned)(nlines < resident ? nlines : resident);
    if ((int)nb > s->divblocks) {
        set_error("ns: divergence partial buffer too small");
        return SH_EINVAL;
    }
    ProfScope ps("ns_fourier_x", s->st);
    hipLaunchKernelGGL((k_ns_fourier_xd<P, NT, SHARE, G3>), dim3(nb, (unsigned)s->nboxes), dim3(64), 0, s->st,
                       (double2*)s->Wf, (double2*)s->V1f, (double2*)s->N0f, (const double2*)tw, s->ny, s->nz, nzp,
                       s->N * s->nboxes, (const StepCoef<double>*)s->coefs, s->divpart, s->N, s->divblocks + 1);
    SH_LAUNCH_CHECK("k_ns_fourier_xd");
    s->div_nb = (int)nb;
    s->w_primed = true;
    s->w5 = false;
    return SH_OK;
}

// kernel E with the merged AB2 history (YS): converts a (V1, N0) state first;
// where the merged form does not apply, the (V1, N0) kernel of the same plan.
// CAP > 0 limits the persistent grid to CAP blocks per compute unit (the
// kernel's LDS admits five one-wave blocks, one more than a CU has SIMDs)
template <typename T, class P, bool NT = true, int CAP = 0, bool E5 = false, bool REV = false, int LAUX = 0>
int launch_fourier_xy(sh_ns_s* s) {
    static_assert(sizeof(T) == 8, "FP64 only");
    // a (V1, N0) state is converted only outside a graph capture
    const bool cap = capturing(s->st);
    if (!y_ok(s) || s->nboxes != 1 || (!s->yform && cap)) return launch_fourier_xd<T, P, NT, true, true>(s);
    // inside a capture the emitted form stays what the captured y pass expects
    const bool e5 = cap ? s->w5 : (E5 && y_curl_ok(s));
    const C2<T>* tw = (const C2<T>*)twiddles(P::N, SH_F64);
    if (!tw) return SH_ENOMEM;
    if (!s->yform) RC((launch_yform<T, true>(s)));
    const int nzp = s->half ? s->nz / 2 : s->nz;
    const long nlines = (long)s->ny * nzp;
    static int resident = 0;
    if (resident == 0) {
        int dev = 0, v = 0, per = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || v <= 0)
            v = 256;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k_ns_fourier_xd<P, NT, true, true, true, E5, REV, LAUX>,
                                                         64, 0) != hipSuccess ||
            per <= 0)
            per = 1;
        if (CAP > 0 && per > CAP) per = CAP;
        resident = v * per;
    }
    const unsigned nb = (unsigned)(nlines < resident ? nlines : resident);
    if ((int)nb > s->divblocks) {
        set_error("ns: divergence partial buffer too small");
        return SH_EINVAL;
    }
    ProfScope ps("ns_fourier_x", s->st);
    if (e5)
        hipLaunchKernelGGL((k_ns_fourier_xd<P, NT, true, true, true, true, REV, LAUX>), dim3(nb, 1u), dim3(64), 0, s->st,
                           (double2*)s->Wf, (double2*)s->V1f, (double2*)s->N0f, (const double2*)tw, s->ny, s->nz,
                           nzp, s->N, (const StepCoef<double>*)s->coefs, s->divpart, s->N, s->divblocks + 1);
    else
        hipLaunchKernelGGL((k_ns_fourier_xd<P, NT, true, true, true, false, REV, LAUX>), dim3(nb, 1u), dim3(64), 0, s->st,
                           (double2*)s->Wf, (double2*)s->V1f, (double2*)s->N0f, (const double2*)tw, s->ny, s->nz,
                           nzp, s->N, (const StepCoef<double>*)s->coefs, s->divpart, s->N, s->divblocks + 1);
    SH_LAUNCH_CHECK("k_ns_fourier_xd");
    s->div_nb = (int)nb;
    s->w_primed = true;
    s->yform = true;
    s->w5 = e5;
    return SH_OK;
}

// FP32 kernel E with LDS-streamed inputs and the merged AB2 history
// (k_ns_fourier_xs): converts a (V1, N0) state first; where the merged form
// does not apply, the register-loaded (V1, N0) kernel of the same plan
template <typename T, class P, bool NT = false, bool E5 = false, bool REV = false, int LAUX = 0>
int launch_fourier_xs(sh_ns_s* s) {
    static_assert(sizeof(T) == 4, "FP32 only");
    // a (V1, N0) state is converted only outside a graph capture
    const bool cap = capturing(s->st);
    if (!y_ok(s) || s->nboxes != 1 || (!s->yform && cap)) return launch_fourier_x<T, P, 1, true, false, NT>(s);
    // inside a capture the emitted form stays what the captured y pass expects
    const bool e5 = cap ? s->w5 : (E5 && y_curl_ok(s));
    const C2<T>* tw = (const C2<T>*)twiddles(P::N, SH_F32);
    if (!tw) return SH_ENOMEM;
    if (!s->yform) RC((launch_yform<T, true>(s)));
    const int nzp = s->half ? s->nz / 2 : s->nz;
    const long nlines = (long)s->ny * nzp;
    static int resident[2] = {0, 0};  // the grid of each emitted form's kernel
    int& res = resident[e5 ? 1 : 0];
    if (res == 0) {
        int dev = 0, v = 0, per = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || v <= 0)
            v = 256;
        const hipError_t e = e5 ? hipOccupancyMaxActiveBlocksPerMultiprocessor(
                                      &per, k_ns_fourier_xs<P, NT, true, REV, LAUX>, 64, 0)
                                : hipOccupancyMaxActiveBlocksPerMultiprocessor(
                                      &per, k_ns_fourier_xs<P, NT, false, REV, LAUX>, 64, 0);
        if (e != hipSuccess || per <= 0) per = 1;
        res = v * per;
    }
    const unsigned nb = (unsigned)(nlines < res ? nlines : res);
    if ((int)nb > s->divblocks) {
        set_error("ns: divergence partial buffer too small");
        return SH_EINVAL;
    }
    ProfScope ps("ns_fourier_x", s->st);
    if (e5)
        hipLaunchKernelGGL((k_ns_fourier_xs<P, NT, true, REV, LAUX>), dim3(nb), dim3(64), 0, s->st,
                           (float2*)s->Wf, (float2*)s->N0f, (const float2*)tw, s->ny, s->nz, nzp, s->N,
                           (const StepCoef<float>*)s->coefs, s->divpart);
    else
        hipLaunchKernelGGL((k_ns_fourier_xs<P, NT, false, REV, LAUX>), dim3(nb), dim3(64), 0, s->st,
                           (float2*)s->Wf, (float2*)s->N0f, (const float2*)tw, s->ny, s->nz, nzp, s->N,
                           (const StepCoef<float>*)s->coefs, s->divpart);
    SH_LAUNCH_CHECK("k_ns_fourier_xs");
    s->div_nb = (int)nb;
    s->w_primed = true;
    s->yform = true;
    s->w5 = e5;
    return SH_OK;
}

// kernel E with the wavefront-exchange transforms and the merged AB2 history
// (k_ns_fourier_xw, 256-point lines): converts a (V1, N0) state first; where
// the merged form does not apply, the (V1, N0) kernel of the same length
template <typename T, bool NT = false, bool E5 = false, bool REV = false, int LAUX = 0>
int launch_fourier_xw(sh_ns_s* s) {
    // a (V1, N0) state is converted only outside a graph capture
    const bool cap = capturing(s->st);
    if (!y_ok(s) || s->nboxes != 1 || (!s->yform && cap)) {
        if constexpr (sizeof(T) == 8)
            return launch_fourier_xd<T, W256, NT, true, true>(s);
        else
            return launch_fourier_x<T, W256, 1, true, false, NT>(s);
    }
    // inside a capture the emitted form stays what the captured y pass expects
    const bool e5 = cap ? s->w5 : (E5 && y_curl_ok(s));
    const C2<T>* tw = (const C2<T>*)twiddles(256, sizeof(T) == 8 ? SH_F64 : SH_F32);
    if (!tw) return SH_ENOMEM;
    if (s->nx != 256) {
        set_error("ns: the wavefront-exchange kernel E needs nx = 256");
        return SH_EINVAL;
    }
    if (!s->yform) RC((launch_yform<T, true>(s)));
    const int nzp = s->half ? s->nz / 2 : s->nz;
    const long nlines = (long)s->ny * nzp;
    static int resident[2] = {0, 0};  // the grid of each emitted form's kernel
    int& res = resident[e5 ? 1 : 0];
    if (res == 0) {
        int dev = 0, v = 0, per = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || v <= 0)
            v = 256;
        const hipError_t e =
            e5 ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k_ns_fourier_xw<T, NT, true, REV, LAUX>, 64, 0)
               : hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k_ns_fourier_xw<T, NT, false, REV, LAUX>, 64, 0);
        if (e != hipSuccess || per <= 0) per = 1;
        res = v * per;
    }
    const unsigned nb = (unsigned)(nlines < res ? nlines : res);
    if ((int)nb > s->divblocks) {
        set_error("ns: divergence partial buffer too small");
        return SH_EINVAL;
    }
    ProfScope ps("ns_fourier_x", s->st);
    if (e5)
        hipLaunchKernelGGL((k_ns_fourier_xw<T, NT, true, REV, LAUX>), dim3(nb), dim3(64), 0, s->st, (C2<T>*)s->Wf,
                           (C2<T>*)s->N0f, tw, s->ny, s->nz, nzp, s->N, (const StepCoef<T>*)s->coefs, s->divpart);
    else
        hipLaunchKernelGGL((k_ns_fourier_xw<T, NT, false, REV, LAUX>), dim3(nb), dim3(64), 0, s->st, (C2<T>*)s->Wf,
                           (C2<T>*)s->N0f, tw, s->ny, s->nz, nzp, s->N, (const StepCoef<T>*)s->coefs, s->divpart);
    SH_LAUNCH_CHECK("k_ns_fourier_xw");
    s->div_nb = (int)nb;
    s->w_primed = true;
    s->yform = true;
    s->w5 = e5;
    return SH_OK;
}

// ---------------------------------------------------------------- tunable fused schedules
// X(n, plan, columns or lines per block); the first entry of a length is its
// default.  Kernel C holds up to seven fields per thread and kernel E three
// products plus the new V, so FP64 defaults to the 4-values-per-field plans.
#define SH_CROSS_F64_VARIANTS(X)                                                                  \
    X(8, P8, 16) X(12, Q12, 16) X(16, P16, 16) X(24, Q24, 16) X(32, Q32, 16) X(48, Q48, 16)       \
    X(64, W64, 16) X(64, W64, 8) X(96, Q96, 16) X(96, Q96, 8) X(128, W128, 16) X(256, W256, 8)    \
    X(256, W256, 4) X(256, Q256, 8) X(256, Q256, 4) X(512, W512, 4) X(512, W512, 2) X(512, P512, 4)
#define SH_CROSS_F32_VARIANTS(X)                                                                  \
    X(8, P8, 16) X(12, Q12, 16) X(16, P16, 16) X(24, Q24, 16) X(32, Q32, 16) X(48, Q48, 16)       \
    X(64, P64, 16) X(64, W64, 16) X(64, P64, 8) X(64, W64, 8)                                     \
    X(96, Q96, 16) X(96, Q96, 8) X(128, Q128, 16) X(256, Q256, 16)                                \
    X(256, Q256, 8) X(256, W256, 16) X(256, W256, 8) X(512, P512, 8) X(512, P512, 4)              \
    X(512, W512, 4)
#define SH_FOURIER_F64_VARIANTS(X)                                                                \
    X(8, P8, 256) X(12, Q12, 64) X(16, P16, 64) X(24, Q24, 32) X(32, Q32, 32) X(48, Q48, 21)      \
    X(64, W64, 16) X(64, P64, 32) X(64, W64, 8) X(64, P64, 8)                                     \
    X(96, Q96, 10) X(128, W128, 8) X(256, W256, 4) X(256, Q256, 8)                                \
    X(256, W256, 2) X(256, W256, 1) X(512, W512, 2) X(512, P512, 4)
#define SH_FOURIER_F32_VARIANTS(X)                                                                \
    X(8, P8, 256) X(12, Q12, 64) X(16, P16, 64) X(24, Q24, 32) X(32, Q32, 32) X(48, Q48, 21)      \
    X(64, P64, 32) X(64, W64, 16) X(64, P64, 16) X(64, P64, 8) X(64, W64, 8) X(64, W64, 4)       \
    X(96, Q96, 10) X(128, Q128, 16) X(256, Q256, 8) X(256, W256, 4)                               \
    X(256, Q256, 4) X(512, P512, 4) X(512, W512, 2)

// kernel C of the full spectrum with its inputs streamed into LDS
// (k_ns_cross_zd with HALF = false, interleaved exchange): X(nz, plan, columns per block)
#define SH_CROSS_F64_LDS_VARIANTS(X) X(128, W128, 8) X(256, W256, 8)
#define SH_CROSS_F32_LDS_VARIANTS(X) X(256, W256, 16)
// kernel C of the half-spectrum layout: X(nz, length-nz/2 plan, columns per block)
#define SH_CROSSH_F64_VARIANTS(X)                                                                 \
    X(16, P8, 16) X(24, Q12, 16) X(32, P16, 16) X(48, Q24, 16) X(64, Q32, 16) X(64, Q32, 8)       \
    X(64, Q32, 32) X(96, Q48, 16)                                                                 \
    X(128, W64, 16) X(128, P64, 16) X(128, W64, 8) X(256, W128, 16) X(256, W128, 8)               \
    X(256, Q128, 16) X(256, Q128, 8) X(512, W256, 8) X(512, W256, 4) X(512, Q256, 8) X(512, Q256, 4)
// kernel C with nontemporal stores of the three products
#define SH_CROSSH_F64_NT_VARIANTS(X) X(256, W128, 16) X(256, W128, 8)
// kernel C with its inputs streamed into LDS (k_ns_cross_zd), plain and
// nontemporal stores: X(nz, length-nz/2 plan, columns per block)
#define SH_CROSSH_F64_LDS_VARIANTS(X) \
    X(64, Q32, 16) X(64, Q32, 8) X(128, W64, 16) X(128, W64, 8) X(256, W128, 16) X(256, W128, 8)
// the same (interleaved exchange) with the packing factors (RT = 1) or also
// the stage twiddles (RT = 2) held in registers: X(nz, plan, columns per block)
#define SH_CROSSH_F64_RT_VARIANTS(X) X(128, W64, 8) X(256, W128, 8) X(256, W128, 16) X(512, W256, 8)
#define SH_CROSSH_F32_NT_VARIANTS(X) X(256, Q128, 16)
#define SH_CROSSH_F32_LDS_VARIANTS(X) X(256, W128, 16)
// FP32 kernel C streamed into LDS with nontemporal loads: X(nz, plan, columns per block)
#define SH_CROSSH_F32_LNT_VARIANTS(X) X(256, W128, 16) X(512, W256, 16)
// FP32 with the stage twiddles and packing factors in registers, nontemporal
// loads (at 256^3: 127-129 us against 139-141; 8 columns per block ran 211-256)
#define SH_CROSSH_F32_RTL_VARIANTS(X) X(256, W128, 16)
// the same for the 64^3 boxes of the ensembles (packed lines of 32; no plan of
// 48 has one butterfly of all a thread's values in its first and last stage)
#define SH_CROSSH_F32_RTL_SMALL_VARIANTS(X) X(64, Q32, 16) X(64, Q32, 32)
// FP64 with the transforms of one LDS exchange (zw_inv, zw_fwd), nontemporal
// loads, plain and nontemporal stores: X(nz, plan, columns per block) (at 256^3:
// 221-222 us against 240; FP32 needs 16 columns per block for whole 128-byte
// rows, which this layout of eight lines per four waves does not give: 212 us
// against 138 with 8)
#define SH_CROSSH_F64_ZW_VARIANTS(X) X(256, W128, 8)
// FP32 with the one-exchange transforms of eight values per thread (zw8_inv,
// zw8_fwd): sixteen columns on four waves, nontemporal loads, plain and
// nontemporal stores: X(nz, plan, columns per block)
#define SH_CROSSH_F32_ZW8_VARIANTS(X) X(256, Q128, 16)
// the same for 256-point packed lines (zs_inv, zs_fwd): sixteen columns on eight waves
#define SH_CROSSH_F32_ZS_VARIANTS(X) X(512, Q256, 16)
#define SH_CROSSH_F32_VARIANTS(X)                                                                 \
    X(16, P8, 16) X(24, Q12, 16) X(32, P16, 16) X(48, Q24, 16) X(64, Q32, 16) X(64, Q32, 8)       \
    X(64, Q32, 32) X(96, Q48, 16)                                                                 \
    X(128, P64, 16) X(128, W64, 16) X(128, P64, 8) X(256, Q128, 16) X(256, W128, 16)              \
    X(256, Q128, 8) X(512, Q256, 16) X(512, Q256, 8) X(512, W256, 16) X(512, W256, 8)

// kernel E with the line's V and N0 stores after the whole update
#define SH_FOURIER_F64_BATCHED_VARIANTS(X) X(64, W64, 16) X(256, W256, 2) X(256, W256, 4) X(512, W512, 2)
#define SH_FOURIER_F32_BATCHED_VARIANTS(X) X(64, W64, 16) X(256, Q256, 8) X(256, W256, 4)
// kernel E with nontemporal stores (V, N0 and the emitted fields)
#define SH_FOURIER_F64_NT_VARIANTS(X) X(256, W256, 2) X(256, W256, 4) X(256, W256, 1)
// kernel E with its inputs streamed into LDS (k_ns_fourier_xd): X(n, plan,
// nontemporal stores, exchanges in the consumed slots, grouped transforms)
#define SH_FOURIER_F64_LDS_VARIANTS(X)                                                        \
    X(256, W256, true, false, false) X(256, W256, false, false, false) X(256, W256, true, true, false) \
    X(256, W256, false, true, false) X(256, W256, true, true, true)
// the same grouped, shared-slot kernel with the merged AB2 history (YS):
// X(n, plan, nontemporal stores, blocks per CU (0: as many as fit))
#define SH_FOURIER_F64_Y_VARIANTS(X) X(256, W256, true, 0) X(256, W256, false, 0) X(256, W256, true, 4)
// the same emitting five fields (E5; the y pass forms the vorticity)
#define SH_FOURIER_F64_Y5_VARIANTS(X) X(256, W256, true, 0) X(256, W256, false, 0)
// the same traversing the lines top plane first (and with nontemporal loads)
#define SH_FOURIER_F64_Y5R_VARIANTS(X) X(256, W256, true, 0) X(256, W256, false, 0)
#define SH_FOURIER_F64_Y5RL_VARIANTS(X) X(256, W256, true, 0) X(256, W256, false, 0)
// the vorticity-forming y pass with all five inputs loaded at once: X(ny,
// plan, columns per block, nontemporal stores, outputs unrolled, twiddles in
// registers); the first is the default at 512 (the not-unrolled forms ran
// 490-540 us at 256^3 against 314-357 us, and were removed)
#define SH_YCURL_F64_VARIANTS(X)                                                                      \
    X(256, Q256, 8, true, true, true) X(256, Q256, 8, true, true, false) X(256, W256, 8, true, true, true) \
    X(256, Q256, 8, false, true, true)                                                                     \
    X(512, W512, 4, true, true, true) X(512, W512, 4, true, true, false) X(512, P512, 4, true, true, true)
#define SH_YCURL_F64_LNT_VARIANTS(X) X(256, Q256, 8, true, true, true) X(256, Q256, 8, false, true, true)
// the staged-load forms: X(ny, plan, columns per block, nontemporal stores,
// twiddles in registers, input fields live at most (2e: loads issued early),
// waves per SIMD); the first is the default at 256.  At 256^3 FP64 (Q256, 8
// columns, two blocks per compute unit) B took 266 us with two fields live
// against 288 us for the unrolled form; at 512 every form ran within 1% of
// the unrolled one (memory-bound), so none is compiled there.
#define SH_ORD_3 1
#define SH_ORD_2 2
#define SH_ORD_2e 3
// (FP32 at 256^3: B 157 -> 145 us with loads issued early, the next kernel C
// 130 -> 137 us, the step within noise; 8-column forms ran 192-206 us)
#define SH_YCURL_F32_ORD_VARIANTS(X) X(256, Q256, 16, true, true, 2e, 2) X(256, Q256, 16, true, true, 2, 2)
#define SH_YCURL_F64_ORD_VARIANTS(X)                                                                  \
    X(256, Q256, 8, true, true, 2, 2) X(256, Q256, 8, false, true, 2, 2) X(256, Q256, 8, true, true, 2e, 2) \
    X(256, Q256, 8, true, true, 3, 2)
// (FP32 at 256: the 16-column form first -- 157 us against 196 for 8 columns --
// because kernel E is tuned before it and picks the five-field form only if
// the default vorticity y pass is fast)
#define SH_YCURL_F32_VARIANTS(X)                                                                      \
    X(256, Q256, 16, true, true, true) X(256, Q256, 8, true, true, true) X(256, W256, 8, true, true, true) \
    X(512, W512, 4, true, true, true) X(512, P512, 4, true, true, true) X(512, P512, 8, true, true, true)
// the register-loaded kernel E with the merged history: X(n, plan, lines per
// block), plain and nontemporal stores
#define SH_FOURIER_F64_YR_VARIANTS(X) X(64, W64, 16) X(96, Q96, 10) X(128, W128, 8) X(512, W512, 2) X(512, P512, 4)
#define SH_FOURIER_F32_YR_VARIANTS(X) \
    X(64, W64, 16) X(64, P64, 32) X(96, Q96, 10) X(128, Q128, 16) X(256, Q256, 8) X(256, W256, 4) X(512, P512, 4) X(512, W512, 2)
// the register-loaded merged kernel E emitting five fields: X(n, plan, lines per block)
#define SH_FOURIER_F64_YR5_VARIANTS(X) X(512, W512, 2) X(512, P512, 4)
#define SH_FOURIER_F32_YR5_VARIANTS(X) X(256, Q256, 8) X(256, W256, 4) X(512, P512, 4) X(512, W512, 2)
// the register-loaded merged kernel E with nontemporal loads: X(n, plan, lines per block)
#define SH_FOURIER_F64_YRL_VARIANTS(X) X(512, W512, 2) X(512, P512, 4)
#define SH_FOURIER_F32_YRL_VARIANTS(X) X(256, Q256, 8) X(256, W256, 4) X(512, P512, 4) X(512, W512, 2)
#define SH_FOURIER_F32_NT_VARIANTS(X) X(256, Q256, 8) X(256, W256, 4)
// the FP32 LDS-streamed kernel E with the merged history (k_ns_fourier_xs):
// X(n, plan, nontemporal stores, five emitted fields, top plane first,
// cache-policy bits of the loads into LDS)
#define SH_FOURIER_F32_S_VARIANTS(X)                                                                    \
    X(256, W256, true, false, true, 2) X(256, W256, false, false, true, 2) X(256, W256, true, false, false, 2) \
    X(256, W256, true, false, true, 0) X(256, W256, false, true, true, 2) X(256, W256, true, true, true, 2) \
    X(512, P512, true, false, true, 2) X(512, P512, false, false, true, 2) X(512, P512, true, true, true, 2)   \
    X(512, P512, false, true, true, 2)
// kernel E with the wavefront-exchange transforms (k_ns_fourier_xw): X(n,
// nontemporal stores, five emitted fields, top plane first, load cache bits)
#define SH_FOURIER_F32_W_VARIANTS(X)                                                                    \
    X(256, true, false, true, 2) X(256, false, false, true, 2) X(256, false, true, true, 2)              \
    X(256, true, true, true, 2)
// (FP64: 331-336 us at 256^3 like the LDS-exchange kernel; the swaps of
// four-dword values cost what the LDS round trips did)
#define SH_FOURIER_F64_W_VARIANTS(X) X(256, false, true, true, 2)

// kernel E without the next step's x-inverse (kernel A runs each step): for
// small grids, where A can overlap other work and E alone is latency-bound
#define SH_FOURIER_F64_NOEMIT_VARIANTS(X) \
    X(32, Q32, 32) X(64, W64, 16) X(64, W64, 8) X(96, Q96, 10) X(256, W256, 2) X(256, W256, 4) X(256, Q256, 8)
#define SH_FOURIER_F32_NOEMIT_VARIANTS(X) \
    X(32, Q32, 32) X(64, P64, 32) X(64, W64, 16) X(64, P64, 8) X(64, W64, 4) X(96, Q96, 10) X(256, Q256, 8)

using NsLaunch = int (*)(sh_ns_s*);
struct NsVariants {
    std::vector<const char*> names;
    std::vector<NsLaunch> fns;
    std::vector<NsLaunch> merged;   // the entries of fns that keep the merged AB2 history
    std::vector<NsLaunch> merged1;  // those that keep it for a single box only
};

template <typename T>
const NsVariants& cross_variants(int n) {
    static const std::map<int, NsVariants> tab = [] {
        std::map<int, NsVariants> m;
#define SH_V(n_, P_, C_)                           \
    m[n_].names.push_back(#P_ " columns=" #C_);    \
    m[n_].fns.push_back(&launch_cross_z<T, P_, C_>);
#define SH_VF(n_, P_, C_)                                                                   \
    m[n_].names.push_back(#P_ " columns=" #C_ " LDS-streamed interleaved exchange");          \
    m[n_].fns.push_back(&launch_cross_zd<T, P_, C_, false, true, false>);                     \
    m[n_].names.push_back(#P_ " columns=" #C_ " LDS-streamed interleaved exchange nontemporal"); \
    m[n_].fns.push_back(&launch_cross_zd<T, P_, C_, true, true, false>);
#define SH_VFR(n_, P_, C_)                                                                             \
    m[n_].names.push_back(#P_ " columns=" #C_ " LDS-streamed interleaved exchange twiddles in registers"); \
    m[n_].fns.push_back(&launch_cross_zd<T, P_, C_, false, true, false, 2>);                              \
    m[n_].names.push_back(#P_ " columns=" #C_ " LDS-streamed interleaved exchange twiddles in registers nontemporal loads"); \
    m[n_].fns.push_back(&launch_cross_zd<T, P_, C_, false, true, false, 2, 2>);                           \
    m[n_].names.push_back(#P_ " columns=" #C_ " LDS-streamed interleaved exchange twiddles in registers nontemporal loads and stores"); \
    m[n_].fns.push_back(&launch_cross_zd<T, P_, C_, true, true, false, 2, 2>);
        if constexpr (sizeof(T) == 8) {
            SH_CROSS_F64_VARIANTS(SH_V)
            SH_CROSS_F64_LDS_VARIANTS(SH_VF)
            SH_CROSS_F64_LDS_VARIANTS(SH_VFR)
        } else {
            SH_CROSS_F32_VARIANTS(SH_V)
            SH_CROSS_F32_LDS_VARIANTS(SH_VF)
        }
#undef SH_VFR
#undef SH_VF
#undef SH_V
        return m;
    }();
    static const NsVariants none;
    auto it = tab.find(n);
    return it == tab.end() ? none : it->second;
}

template <typename T>
const NsVariants& crossh_variants(int n) {
    static const std::map<int, NsVariants> tab = [] {
        std::map<int, NsVariants> m;
#define SH_V(n_, P_, C_)                                \
    m[n_].names.push_back(#P_ " half columns=" #C_);    \
    m[n_].fns.push_back(&launch_cross_zh<T, P_, C_>);
#define SH_VT(n_, P_, C_)                                               \
    m[n_].names.push_back(#P_ " half columns=" #C_ " nontemporal stores");  \
    m[n_].fns.push_back(&launch_cross_zh<T, P_, C_, true>);
#define SH_VD(n_, P_, C_)                                                        \
    m[n_].names.push_back(#P_ " half columns=" #C_ " LDS-streamed");                  \
    m[n_].fns.push_back(&launch_cross_zd<T, P_, C_>);                                 \
    m[n_].names.push_back(#P_ " half columns=" #C_ " LDS-streamed nontemporal stores"); \
    m[n_].fns.push_back(&launch_cross_zd<T, P_, C_, true>);                           \
    m[n_].names.push_back(#P_ " half columns=" #C_ " LDS-streamed interleaved exchange"); \
    m[n_].fns.push_back(&launch_cross_zd<T, P_, C_, false, true>);
#define SH_VDL(n_, P_, C_)                                                                                  \
    m[n_].names.push_back(#P_ " half columns=" #C_ " LDS-streamed interleaved exchange nontemporal loads");           \
    m[n_].fns.push_back(&launch_cross_zd<T, P_, C_, false, true, true, 0, 2>);                                  \
    m[n_].names.push_back(#P_ " half columns=" #C_ " LDS-streamed interleaved exchange nontemporal loads and stores"); \
    m[n_].fns.push_back(&launch_cross_zd<T, P_, C_, true, true, true, 0, 2>);
#define SH_VRL(n_, P_, C_)                                                                                  \
    m[n_].names.push_back(#P_ " half columns=" #C_ " LDS-streamed interleaved exchange twiddles in registers nontemporal loads"); \
    m[n_].fns.push_back(&launch_cross_zd<T, P_, C_, false, true, true, 2, 2>);                                 \
    m[n_].names.push_back(#P_ " half columns=" #C_ " LDS-streamed interleaved exchange twiddles in registers nontemporal loads and stores"); \
    m[n_].fns.push_back(&launch_cross_zd<T, P_, C_, true, true, true, 2, 2>);
#define SH_VZW8(n_, P_, C_)                                                                                 \
    m[n_].names.push_back(#P_ " half columns=" #C_ " LDS-streamed one-exchange transforms of eight values nontemporal loads"); \
    m[n_].fns.push_back(&launch_cross_zd<T, P_, C_, false, true, true, 4, 2>);                                \
    m[n_].names.push_back(#P_ " half columns=" #C_ " LDS-streamed one-exchange transforms of eight values nontemporal loads and stores"); \
    m[n_].fns.push_back(&launch_cross_zd<T, P_, C_, true, true, true, 4, 2>);
#define SH_VZS(n_, P_, C_)                                                                                  \
    m[n_].names.push_back(#P_ " half columns=" #C_ " LDS-streamed one-exchange transforms of eight values nontemporal loads"); \
    m[n_].fns.push_back(&launch_cross_zd<T, P_, C_, false, true, true, 5, 2>);                                \
    m[n_].names.push_back(#P_ " half columns=" #C_ " LDS-streamed one-exchange transforms of eight values nontemporal loads and stores"); \
    m[n_].fns.push_back(&launch_cross_zd<T, P_, C_, true, true, true, 5, 2>);
#define SH_VZW(n_, P_, C_)                                                                                  \
    m[n_].names.push_back(#P_ " half columns=" #C_ " LDS-streamed one-exchange transforms nontemporal loads");     \
    m[n_].fns.push_back(&launch_cross_zd<T, P_, C_, false, true, true, 3, 2>);                                 \
    m[n_].names.push_back(#P_ " half columns=" #C_ " LDS-streamed one-exchange transforms nontemporal loads and stores"); \
    m[n_].fns.push_back(&launch_cross_zd<T, P_, C_, true, true, true, 3, 2>);
#define SH_VR(n_, P_, C_)                                                                                   \
    m[n_].names.push_back(#P_ " half columns=" #C_ " LDS-streamed interleaved exchange packing factors in registers"); \
    m[n_].fns.push_back(&launch_cross_zd<T, P_, C_, false, true, true, 1>);                                    \
    m[n_].names.push_back(#P_ " half columns=" #C_ " LDS-streamed interleaved exchange twiddles in registers");    \
    m[n_].fns.push_back(&launch_cross_zd<T, P_, C_, false, true, true, 2>);                                    \
    m[n_].names.push_back(#P_ " half columns=" #C_ " LDS-streamed interleaved exchange twiddles in registers nontemporal"); \
    m[n_].fns.push_back(&launch_cross_zd<T, P_, C_, true, true, true, 2>);                                     \
    m[n_].names.push_back(#P_ " half columns=" #C_ " LDS-streamed interleaved exchange twiddles in registers nontemporal loads"); \
    m[n_].fns.push_back(&launch_cross_zd<T, P_, C_, false, true, true, 2, 2>);                                 \
    m[n_].names.push_back(#P_ " half columns=" #C_ " LDS-streamed interleaved exchange twiddles in registers nontemporal loads and stores"); \
    m[n_].fns.push_back(&launch_cross_zd<T, P_, C_, true, true, true, 2, 2>);
        if constexpr (sizeof(T) == 8) {
            SH_CROSSH_F64_VARIANTS(SH_V)
            SH_CROSSH_F64_NT_VARIANTS(SH_VT)
            SH_CROSSH_F64_LDS_VARIANTS(SH_VD)
            SH_CROSSH_F64_RT_VARIANTS(SH_VR)
            SH_CROSSH_F64_ZW_VARIANTS(SH_VZW)
        } else {
            SH_CROSSH_F32_VARIANTS(SH_V)
            SH_CROSSH_F32_NT_VARIANTS(SH_VT)
            SH_CROSSH_F32_LDS_VARIANTS(SH_VD)
            SH_CROSSH_F32_LNT_VARIANTS(SH_VDL)
            SH_CROSSH_F32_RTL_VARIANTS(SH_VRL)
            SH_CROSSH_F32_RTL_SMALL_VARIANTS(SH_VRL)
            SH_CROSSH_F32_ZW8_VARIANTS(SH_VZW8)
            SH_CROSSH_F32_ZS_VARIANTS(SH_VZS)
        }
#undef SH_VZS
#undef SH_VZW8
#undef SH_VZW
#undef SH_VRL
#undef SH_VR
#undef SH_VDL
#undef SH_VD
#undef SH_VT
#undef SH_V
        return m;
    }();
    static const NsVariants none;
    auto it = tab.find(n);
    return it == tab.end() ? none : it->second;
}

// the half-spectrum layout needs an even nz whose half has a fused plan
bool half_supported(int prec, int nz) {
    return nz % 2 == 0 && (prec == SH_F64 ? !crossh_variants<double>(nz).fns.empty()
                                          : !crossh_variants<float>(nz).fns.empty());
}

template <typename T>
const NsVariants& fourier_variants(int n) {
    static const std::map<int, NsVariants> tab = [] {
        std::map<int, NsVariants> m;
#define SH_V(n_, P_, L_)                         \
    m[n_].names.push_back(#P_ " lines=" #L_);    \
    m[n_].fns.push_back(&launch_fourier_x<T, P_, L_>);
#define SH_VB(n_, P_, L_)                                       \
    m[n_].names.push_back(#P_ " lines=" #L_ " deferred stores"); \
    m[n_].fns.push_back(&launch_fourier_x<T, P_, L_, true, true>);
#define SH_VN(n_, P_, L_)                                  \
    m[n_].names.push_back(#P_ " lines=" #L_ " no-emit");   \
    m[n_].fns.push_back(&launch_fourier_x<T, P_, L_, false>);
#define SH_VT(n_, P_, L_)                                             \
    m[n_].names.push_back(#P_ " lines=" #L_ " nontemporal stores");   \
    m[n_].fns.push_back(&launch_fourier_x<T, P_, L_, true, false, true>);
#define SH_VS(n_, P_, NT_, SH_, G3_)                                                               \
    m[n_].names.push_back(#P_ " LDS-streamed persistent lines, nontemporal=" #NT_ ", shared exchange=" #SH_ \
                          ", grouped=" #G3_);                                                          \
    m[n_].fns.push_back(&launch_fourier_xd<T, P_, NT_, SH_, G3_>);
#define SH_VY(n_, P_, NT_, CAP_)                                                                              \
    m[n_].names.push_back(#P_ " LDS-streamed persistent lines, merged AB2 history, nontemporal=" #NT_          \
                          ", blocks per CU=" #CAP_);                                                            \
    m[n_].fns.push_back(&launch_fourier_xy<T, P_, NT_, CAP_>);                                                 \
    m[n_].merged1.push_back(&launch_fourier_xy<T, P_, NT_, CAP_>);
#define SH_VY5(n_, P_, NT_, CAP_)                                                                             \
    m[n_].names.push_back(#P_ " LDS-streamed persistent lines, merged AB2 history, five emitted fields, "      \
                          "nontemporal=" #NT_);                                                                 \
    m[n_].fns.push_back(&launch_fourier_xy<T, P_, NT_, CAP_, true>);                                           \
    m[n_].merged1.push_back(&launch_fourier_xy<T, P_, NT_, CAP_, true>);
#define SH_VY5R(n_, P_, NT_, CAP_)                                                                            \
    m[n_].names.push_back(#P_ " LDS-streamed persistent lines, merged AB2 history, five emitted fields, "      \
                          "nontemporal=" #NT_ ", top plane first");                                            \
    m[n_].fns.push_back(&launch_fourier_xy<T, P_, NT_, CAP_, true, true>);                                     \
    m[n_].merged1.push_back(&launch_fourier_xy<T, P_, NT_, CAP_, true, true>);
#define SH_VY5RL(n_, P_, NT_, CAP_)                                                                           \
    m[n_].names.push_back(#P_ " LDS-streamed persistent lines, merged AB2 history, five emitted fields, "      \
                          "nontemporal=" #NT_ ", top plane first, nontemporal loads");                         \
    m[n_].fns.push_back(&launch_fourier_xy<T, P_, NT_, CAP_, true, true, 2>);                                  \
    m[n_].merged1.push_back(&launch_fourier_xy<T, P_, NT_, CAP_, true, true, 2>);
#define SH_VM5(n_, P_, L_)                                                                      \
    m[n_].names.push_back(#P_ " lines=" #L_ " merged AB2 history, five emitted fields, nontemporal stores"); \
    m[n_].fns.push_back(&launch_fourier_xm<T, P_, L_, false, true, true>);                                \
    m[n_].merged.push_back(&launch_fourier_xm<T, P_, L_, false, true, true>);
#define SH_VML(n_, P_, L_)                                                                                  \
    m[n_].names.push_back(#P_ " lines=" #L_ " merged AB2 history, nontemporal loads");                        \
    m[n_].fns.push_back(&launch_fourier_xm<T, P_, L_, false, false, false, true>);                           \
    m[n_].merged.push_back(&launch_fourier_xm<T, P_, L_, false, false, false, true>);                        \
    m[n_].names.push_back(#P_ " lines=" #L_ " merged AB2 history, nontemporal loads and stores");             \
    m[n_].fns.push_back(&launch_fourier_xm<T, P_, L_, false, true, false, true>);                            \
    m[n_].merged.push_back(&launch_fourier_xm<T, P_, L_, false, true, false, true>);
#define SH_VXS(n_, P_, NT_, E5_, REV_, LAUX_)                                                            \
    m[n_].names.push_back(E5_ ? #P_ " LDS-streamed persistent lines, merged AB2 history, five emitted fields, "  \
                                "nontemporal=" #NT_ ", top plane first=" #REV_ ", load cache bits=" #LAUX_      \
                              : #P_ " LDS-streamed persistent lines, merged AB2 history, six emitted fields, "   \
                                "nontemporal=" #NT_ ", top plane first=" #REV_ ", load cache bits=" #LAUX_);     \
    m[n_].fns.push_back(&launch_fourier_xs<T, P_, NT_, E5_, REV_, LAUX_>);                                  \
    m[n_].merged1.push_back(&launch_fourier_xs<T, P_, NT_, E5_, REV_, LAUX_>);
#define SH_VXW(n_, NT_, E5_, REV_, LAUX_)                                                                   \
    m[n_].names.push_back(E5_ ? "W256 LDS-streamed persistent lines, wavefront exchanges, merged AB2 history, five "   \
                                "emitted fields, nontemporal=" #NT_ ", top plane first=" #REV_ ", load cache bits=" #LAUX_ \
                              : "W256 LDS-streamed persistent lines, wavefront exchanges, merged AB2 history, six "    \
                                "emitted fields, nontemporal=" #NT_ ", top plane first=" #REV_ ", load cache bits=" #LAUX_); \
    m[n_].fns.push_back(&launch_fourier_xw<T, NT_, E5_, REV_, LAUX_>);                                       \
    m[n_].merged1.push_back(&launch_fourier_xw<T, NT_, E5_, REV_, LAUX_>);
#define SH_VM(n_, P_, L_)                                                                     \
    m[n_].names.push_back(#P_ " lines=" #L_ " merged AB2 history");                            \
    m[n_].fns.push_back(&launch_fourier_xm<T, P_, L_>);                                        \
    m[n_].merged.push_back(&launch_fourier_xm<T, P_, L_>);                                     \
    m[n_].names.push_back(#P_ " lines=" #L_ " merged AB2 history nontemporal stores");         \
    m[n_].fns.push_back(&launch_fourier_xm<T, P_, L_, false, true>);                           \
    m[n_].merged.push_back(&launch_fourier_xm<T, P_, L_, false, true>);
        if constexpr (sizeof(T) == 8) {
            SH_FOURIER_F64_VARIANTS(SH_V)
            SH_FOURIER_F64_NOEMIT_VARIANTS(SH_VN)
            SH_FOURIER_F64_BATCHED_VARIANTS(SH_VB)
            SH_FOURIER_F64_NT_VARIANTS(SH_VT)
            SH_FOURIER_F64_LDS_VARIANTS(SH_VS)
            SH_FOURIER_F64_Y_VARIANTS(SH_VY)
            SH_FOURIER_F64_Y5_VARIANTS(SH_VY5)
            SH_FOURIER_F64_Y5R_VARIANTS(SH_VY5R)
            SH_FOURIER_F64_Y5RL_VARIANTS(SH_VY5RL)
            SH_FOURIER_F64_YR_VARIANTS(SH_VM)
            SH_FOURIER_F64_YR5_VARIANTS(SH_VM5)
            SH_FOURIER_F64_YRL_VARIANTS(SH_VML)
            SH_FOURIER_F64_W_VARIANTS(SH_VXW)
        } else {
            SH_FOURIER_F32_VARIANTS(SH_V)
            SH_FOURIER_F32_NOEMIT_VARIANTS(SH_VN)
            SH_FOURIER_F32_BATCHED_VARIANTS(SH_VB)
            SH_FOURIER_F32_NT_VARIANTS(SH_VT)
            SH_FOURIER_F32_YR_VARIANTS(SH_VM)
            SH_FOURIER_F32_YR5_VARIANTS(SH_VM5)
            SH_FOURIER_F32_YRL_VARIANTS(SH_VML)
            SH_FOURIER_F32_S_VARIANTS(SH_VXS)
            SH_FOURIER_F32_W_VARIANTS(SH_VXW)
        }
#undef SH_VXW
#undef SH_VXS
#undef SH_VML
#undef SH_VM5
#undef SH_VM
#undef SH_VY5
#undef SH_VY5R
#undef SH_VY5RL
#undef SH_VY
#undef SH_VS
#undef SH_VT
#undef SH_V
#undef SH_VN
#undef SH_VB
        return m;
    }();
    static const NsVariants none;
    auto it = tab.find(n);
    return it == tab.end() ? none : it->second;
}

template <typename T>
int run_variant(const NsVariants& vl, sh_ns_s* s, int n, int role) {
    if (vl.fns.empty()) {
        set_error("ns: unsupported grid size");
        return SH_EUNSUPPORTED;
    }
    int v = tune_choice(sizeof(T) == 8 ? SH_F64 : SH_F32, n, role);
    if (v < 0 || v >= (int)vl.fns.size()) v = 0;
    return vl.fns[v](s);
}

template <typename T>
int dispatch_inv_x(sh_ns_s* s) {
    switch (s->nx) {
#define SH_CASE(n_, P_, C_) \
    case n_:                \
        return launch_inv_x<T, P_>(s);
        SH_FOR_EACH_FUSED_SIZE(SH_CASE)
#undef SH_CASE
    }
    return SH_EUNSUPPORTED;
}
template <typename T>
int dispatch_cross_z(sh_ns_s* s) {
    if (s->half) return run_variant<T>(crossh_variants<T>(s->nz), s, s->nz, ROLE_CROSS_H);
    return run_variant<T>(cross_variants<T>(s->nz), s, s->nz, ROLE_CROSS);
}

template <typename T>
int dispatch_fourier_x(sh_ns_s* s) {
    return run_variant<T>(fourier_variants<T>(s->nx), s, s->nx, ROLE_FOURIER);
}

template <typename T>
const NsVariants& ycurl_variants(int n) {
    static const std::map<int, NsVariants> tab = [] {
        std::map<int, NsVariants> m;
#define SH_V(n_, P_, C_, NT_, U_, RT_)                                                              \
    m[n_].names.push_back(#P_ " columns=" #C_ " vorticity on load, nontemporal=" #NT_ ", unrolled=" #U_   \
                          ", twiddles in registers=" #RT_);                                              \
    m[n_].fns.push_back(&launch_y_curl<T, P_, C_, NT_, U_, RT_>);
#define SH_VL(n_, P_, C_, NT_, U_, RT_)                                                             \
    m[n_].names.push_back(#P_ " columns=" #C_ " vorticity on load, nontemporal=" #NT_ ", unrolled=" #U_   \
                          ", twiddles in registers=" #RT_ ", nontemporal loads");                        \
    m[n_].fns.push_back(&launch_y_curl<T, P_, C_, NT_, U_, RT_, true>);
#define SH_VO(n_, P_, C_, NT_, RT_, LIVE_, OCC_)                                                         \
    m[n_].names.push_back(#P_ " columns=" #C_ " vorticity on load, staged loads (" #LIVE_ " fields live), nontemporal=" #NT_ \
                          ", twiddles in registers=" #RT_ ", waves per SIMD=" #OCC_);                    \
    m[n_].fns.push_back(&launch_y_curl<T, P_, C_, NT_, true, RT_, false, SH_ORD_##LIVE_, OCC_>);
        if constexpr (sizeof(T) == 8) {
            SH_YCURL_F64_ORD_VARIANTS(SH_VO)
            SH_YCURL_F64_VARIANTS(SH_V)
            SH_YCURL_F64_LNT_VARIANTS(SH_VL)
        } else {
            SH_YCURL_F32_VARIANTS(SH_V)
            SH_YCURL_F32_ORD_VARIANTS(SH_VO)
        }
#undef SH_VO
#undef SH_VL
#undef SH_V
        return m;
    }();
    static const NsVariants none;
    auto it = tab.find(n);
    return it == tab.end() ? none : it->second;
}

template <typename T>
int dispatch_y_curl(sh_ns_s* s) {
    return run_variant<T>(ycurl_variants<T>(s->ny), s, s->ny, ROLE_YCURL);
}

bool y_curl_ok(sh_ns_s* s) {
    return s->nboxes == 1 && !(s->prec == SH_F64 ? ycurl_variants<double>(s->ny) : ycurl_variants<float>(s->ny))
                                 .fns.empty();
}

// whether this step's kernel E keeps the merged AB2 history
bool fourier_merged(sh_ns_s* s) {
    if (!y_ok(s)) return false;
    const NsVariants& vl = s->prec == SH_F64 ? fourier_variants<double>(s->nx) : fourier_variants<float>(s->nx);
    if (vl.fns.empty()) return false;
    int v = tune_choice(s->prec, s->nx, ROLE_FOURIER);
    if (v < 0 || v >= (int)vl.fns.size()) v = 0;
    for (NsLaunch f : vl.merged)
        if (f == vl.fns[v]) return true;
    if (s->nboxes == 1)
        for (NsLaunch f : vl.merged1)
            if (f == vl.fns[v]) return true;
    return false;
}

// kernels B, C, D over the columns [x0, x0 + ncx) on stream st
template <typename T>
int step_bcd(sh_ns_s* s, hipStream_t st, int x0, int ncx) {
    const long sxy = (long)s->nx * s->ny;
    const int nzp = s->half ? s->nz / 2 : s->nz;  // planes the y passes cover
    Epi<T> plain{T(1), 0, 1, -1, -1};
    C2<T>* w = (C2<T>*)s->Wf + x0;
    s->cst = st;
    s->cx0 = x0;
    s->cnx = ncx;
    if (s->w5) {
        RC(dispatch_y_curl<T>(s));  // its own profiling scope, "ns_y_inv"
    } else {
        ProfScope ps("ns_y_inv", st);
        RC(pass_strided<T>(w, s->ny, ncx, 1, nzp, sxy, s->nx, 6 * s->nboxes, s->N, +1, plain, st, ROLE_Y));
    }
    RC(dispatch_cross_z<T>(s));
    {
        ProfScope ps("ns_y_fwd", st);
        RC(pass_strided<T>(w, s->ny, ncx, 1, nzp, sxy, s->nx, 3 * s->nboxes, s->N, -1, plain, st, ROLE_Y));
    }
    return SH_OK;
}

// kernels A (after a restart), B, C, D of a step: everything before the
// Fourier update, which is the only kernel that reads the body force.  With
// overlap, B-D run as two independent column halves on two streams, so one
// half's transform passes overlap the other half's fused z kernel; kernel E
// waits for both.
template <typename T>
int step_fused_front(sh_ns_s* s) {
    if (!s->nonlinear) {
        set_error("ns: the fused schedule always includes the nonlinear term");
        return SH_EINVAL;
    }
    // a merged-history state goes back to (V1, N0) before a kernel E that
    // needs them (while W still holds the x-inverse of V1)
    if (s->yform && !fourier_merged(s)) RC(ns_canonical(s));
    // kernel A runs once after a (re)start; afterwards kernel E emits its output
    if (!s->w_primed) {
        RC(dispatch_inv_x<T>(s));
        s->w_primed = true;
    }
    if (!s->overlap || s->nx < 2) return step_bcd<T>(s, s->st, 0, s->nx);
    const int h = s->nx / 2;
    SH_HIP(hipEventRecord(s->ev_fork, s->st));
    SH_HIP(hipStreamWaitEvent(s->st2, s->ev_fork, 0));
    RC(step_bcd<T>(s, s->st2, h, s->nx - h));
    RC(step_bcd<T>(s, s->st, 0, h));
    SH_HIP(hipEventRecord(s->ev_join, s->st2));
    SH_HIP(hipStreamWaitEvent(s->st, s->ev_join, 0));
    return SH_OK;
}

template <typename T>
int step_fused(sh_ns_s* s) {
    RC(step_fused_front<T>(s));
    RC(dispatch_fourier_x<T>(s));
    s->have_div = true;
    return SH_OK;
}

// time `reps` steps (after one warm step) of the solver as it is now configured
// step_cb: the caller's step (e.g. one that adds work on other streams and
// joins them back into s->st); null: the solver's own fused step
int time_steps(sh_ns_s* s, int reps, hipEvent_t e0, hipEvent_t e1, double* ms, int (*step_cb)(void*),
               void* ctx) {
    auto step = [&]() {
        if (step_cb) return step_cb(ctx);
        return s->prec == SH_F64 ? step_fused<double>(s) : step_fused<float>(s);
    };
    RC(step());
    SH_HIP(hipEventRecord(e0, s->st));
    for (int r = 0; r < reps; r++) RC(step());
    SH_HIP(hipEventRecord(e1, s->st));
    SH_HIP(hipEventSynchronize(e1));
    float t = 0;
    SH_HIP(hipEventElapsedTime(&t, e0, e1));
    *ms = t / reps;
    return SH_OK;
}

}  // namespace

namespace sh {
int ns_autotune_with(sh_ns_t s, int reps, int (*step_cb)(void*), void* ctx);

int ns_variant_count(int precision, int n, int role) {
    if (role == ROLE_YCURL)
        return precision == SH_F64 ? (int)ycurl_variants<double>(n).fns.size()
                                   : (int)ycurl_variants<float>(n).fns.size();
    if (role == ROLE_CROSS_H)
        return precision == SH_F64 ? (int)crossh_variants<double>(n).fns.size()
                                   : (int)crossh_variants<float>(n).fns.size();
    if (role == ROLE_CROSS)
        return precision == SH_F64 ? (int)cross_variants<double>(n).fns.size()
                                   : (int)cross_variants<float>(n).fns.size();
    return precision == SH_F64 ? (int)fourier_variants<double>(n).fns.size()
                               : (int)fourier_variants<float>(n).fns.size();
}

const char* ns_variant_name(int precision, int n, int role, int v) {
    const NsVariants& vl =
        role == ROLE_YCURL   ? (precision == SH_F64 ? ycurl_variants<double>(n) : ycurl_variants<float>(n))
        : role == ROLE_CROSS_H ? (precision == SH_F64 ? crossh_variants<double>(n) : crossh_variants<float>(n))
        : role == ROLE_CROSS ? (precision == SH_F64 ? cross_variants<double>(n) : cross_variants<float>(n))
                             : (precision == SH_F64 ? fourier_variants<double>(n) : fourier_variants<float>(n));
    return v >= 0 && v < (int)vl.names.size() ? vl.names[v] : "default";
}
}  // namespace sh

extern "C" {

// Times every schedule of the solver's tunable kernels (y pass, fused z,
// fused Fourier/x) on this solver's grid, keeps the fastest of each, and
// restores the solver state afterwards.  Synchronous.
int sh_ns_autotune(sh_ns_t s, int reps) { return ns_autotune_with(s, reps, nullptr, nullptr); }

}  // extern "C"

namespace sh {
// the solver's schedule tuner timing `step_cb(ctx)` steps (null: its own step);
// the solver state is saved and restored around the timing
int ns_autotune_with(sh_ns_t s, int reps, int (*step_cb)(void*), void* ctx) {
    if (!s || reps <= 0) return SH_EINVAL;
    if (!s->fused) {
        set_error("sh_ns_autotune: the op-by-op schedule has nothing to tune");
        return SH_EINVAL;
    }
    const size_t e = s->prec == SH_F64 ? sizeof(double2) : sizeof(float2);
    const size_t fb = e * (size_t)s->N * s->nboxes;  // one field of every box
    void* save = nullptr;
    RC(pool_alloc(&save, 12 * fb));
    char* sv = (char*)save;
    SH_HIP(hipMemcpyAsync(sv, s->V1f, 3 * fb, hipMemcpyDeviceToDevice, s->st));
    SH_HIP(hipMemcpyAsync(sv + 3 * fb, s->N0f, 3 * fb, hipMemcpyDeviceToDevice, s->st));
    SH_HIP(hipMemcpyAsync(sv + 6 * fb, s->Wf, 6 * fb, hipMemcpyDeviceToDevice, s->st));
    const bool primed = s->w_primed, yform = s->yform, w5 = s->w5;
    hipEvent_t e0, e1;
    SH_HIP(hipEventCreate(&e0));
    SH_HIP(hipEventCreate(&e1));
    int rc = SH_OK;
    // the vorticity-forming y pass last: it runs only after a kernel E that
    // emits five fields, so it is timed with the chosen E
    // Two rounds over the roles: a kernel's stores (plain or nontemporal) and
    // its traversal order decide how much of its output the next kernel finds
    // in the last-level cache, so each role is timed in whole steps and chosen
    // again once the others have their picks.
    const int roles[4] = {ROLE_Y, s->half ? ROLE_CROSS_H : ROLE_CROSS, ROLE_FOURIER, ROLE_YCURL};
    const int lens[4] = {s->ny, s->nz, s->nx, s->ny};
    {
        // warm steps first (their state is restored with the rest): the
        // device's clocks settle before the schedules are timed
        double ms = 0;
        rc = time_steps(s, 8 * reps, e0, e1, &ms, step_cb, ctx);
    }
    for (int k = 0; k < 8 && !rc; k++) {
        const int role = roles[k % 4], n = lens[k % 4];
        const int nv = role == ROLE_Y ? pass_variant_count(s->prec, n, role) : ns_variant_count(s->prec, n, role);
        if (nv <= 1) continue;
        std::vector<double> t(nv, INFINITY);
        auto time_variant = [&](int v, int k) {
            tune_set(s->prec, n, role, v, 0);
            double ms = 0;
            rc = time_steps(s, k, e0, e1, &ms, step_cb, ctx);
            if (!rc) t[v] = std::min(t[v], ms);
        };
        // two sweeps, then a playoff: the three fastest timed alternately three
        // more times over three times the steps, the least time of each
        // counting (schedules within a percent of each other are otherwise
        // told apart by the noise of one short timing)
        for (int pass = 0; pass < 2; pass++)
            for (int v = 0; v < nv && !rc; v++) time_variant(v, reps);
        std::vector<int> top(nv);
        for (int v = 0; v < nv; v++) top[v] = v;
        std::sort(top.begin(), top.end(), [&](int a, int b) { return t[a] < t[b]; });
        top.resize(std::min(nv, 3));
        for (int v : top) t[v] = INFINITY;  // the finalists are judged by their long timings only
        for (int r = 0; r < 3 && !rc; r++)
            for (int v : top)
                if (!rc) time_variant(v, 3 * reps);
        int best = top[0];
        for (int v : top)
            if (t[v] < t[best]) best = v;
        if (!rc) tune_set(s->prec, n, role, best, t[best]);
    }
    // then the serial or the overlapped (two column halves on two streams) B-D
    if (!rc) {
        double ms[2] = {INFINITY, INFINITY};
        for (int ov = 0; ov < 2 && !rc; ov++) {
            s->overlap = ov;
            rc = time_steps(s, reps, e0, e1, &ms[ov], step_cb, ctx);
        }
        s->overlap = (!rc && ms[1] < ms[0]) ? 1 : 0;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    SH_HIP(hipMemcpyAsync(s->V1f, sv, 3 * fb, hipMemcpyDeviceToDevice, s->st));
    SH_HIP(hipMemcpyAsync(s->N0f, sv + 3 * fb, 3 * fb, hipMemcpyDeviceToDevice, s->st));
    SH_HIP(hipMemcpyAsync(s->Wf, sv + 6 * fb, 6 * fb, hipMemcpyDeviceToDevice, s->st));
    SH_HIP(hipStreamSynchronize(s->st));
    pool_free(save);
    s->w_primed = primed;
    s->yform = yform;
    s->w5 = w5;
    s->have_div = false;
    return rc;
}
}  // namespace sh

namespace {

// per-box step coefficients on the device (exec/ns.cpp:171, 174, 186)
template <typename T>
int upload_coefs_t(sh_ns_s* s) {
    std::vector<StepCoef<T>> h(s->nboxes);
    for (int b = 0; b < s->nboxes; b++) {
        const double nu = s->nus[b];
        const double c1 = s->scheme ? nu * s->K : nu * s->K / 2;
        h[b].Kh = T(s->K / 2);
        h[b].c1 = T(c1);
        h[b].c1d = c1;
        h[b].be = s->scheme;
        h[b].inv_n3 = T(1.0 / ((double)s->nx * s->ny * s->nz));
        h[b].dealias = s->dealias;
    }
    SH_HIP(hipMemcpyAsync(s->coefs, h.data(), sizeof(StepCoef<T>) * s->nboxes, hipMemcpyHostToDevice, s->st));
    SH_HIP(hipStreamSynchronize(s->st));
    return SH_OK;
}

int upload_coefs(sh_ns_s* s) { return s->prec == SH_F64 ? upload_coefs_t<double>(s) : upload_coefs_t<float>(s); }

// complete `count` half-spectrum fields (stride N) to the full spectrum in place
template <typename T>
int expand_half(sh_ns_s* s, void* f, int count) {
    const long nup = (long)s->nx * s->ny * (s->nz - s->nz / 2) * count;
    hipLaunchKernelGGL(k_expand_half<T>, dim3(egrid(nup)), dim3(256), 0, s->st, (C2<T>*)f, count, s->N, s->nx,
                       s->ny, s->nz);
    SH_LAUNCH_CHECK("k_expand_half");
    return SH_OK;
}

int expand_state(sh_ns_s* s) {
    RC(ns_canonical(s));
    const int nf = 3 * s->nboxes;
    if (s->prec == SH_F64) {
        RC(expand_half<double>(s, s->V1f, nf));
        return expand_half<double>(s, s->N0f, nf);
    }
    RC(expand_half<float>(s, s->V1f, nf));
    return expand_half<float>(s, s->N0f, nf);
}

// largest |f(k) - conj f(-k)| and |f(k)| over 3 FP64 fields
int herm_dev(sh_ns_s* s, const double2* f, double* dev, double* mag) {
    constexpr int g = 1024;
    double* part = nullptr;
    RC(pool_alloc((void**)&part, sizeof(double) * 2 * g));
    hipLaunchKernelGGL(k_herm_dev, dim3(g), dim3(256), 0, s->st, f, 3, s->N, s->nx, s->ny, s->nz, part);
    std::vector<double> h(2 * g);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(h.data(), part, sizeof(double) * 2 * g, hipMemcpyDeviceToHost, s->st);
    if (e == hipSuccess) e = hipStreamSynchronize(s->st);
    pool_free(part);
    if (e != hipSuccess) return hip_fail(e, "herm_dev");
    double d = 0, m = 0;
    for (int b = 0; b < g; b++) {
        if (!(h[2 * b] <= d)) d = h[2 * b];  // NaN propagates
        if (!(h[2 * b + 1] <= m)) m = h[2 * b + 1];
    }
    *dev = d;
    *mag = m;
    return SH_OK;
}

// Choose the storage of the fused state after box b's start-up: the half
// layout needs the layout switch on, a supported nz, and every box's start
// (V1, cross0) Hermitian -- V1 to 1e-12 of its size, cross0 to 1e-12 of its
// size or with K*|deviation| below 1e-12 of V1's size (its step contribution).
int decide_layout(sh_ns_s* s, int b) {
    if (!s->fused) return SH_OK;
    bool herm = false;
    if (s->layout == 1 && half_supported(s->prec, s->nz)) {
        double dv, fv, dc, fc;
        RC(herm_dev(s, s->V1, &dv, &fv));
        RC(herm_dev(s, s->cr0, &dc, &fc));
        const double tol = 1e-12;
        herm = dv <= tol * fv && (dc <= tol * fc || s->K * dc <= tol * fv);
    }
    s->box_herm[b] = herm;
    bool all = true;
    for (char h : s->box_herm) all = all && h;
    if (s->half && !all) RC(expand_state(s));
    s->half = all;
    s->w_primed = false;
    // the half-layout z kernel's twiddle tables, built now so that a step
    // captured into a graph allocates nothing
    if (s->half && (!twiddles(s->nz / 2, s->prec) || !twiddles(s->nz, s->prec))) return SH_ENOMEM;
    return SH_OK;
}

// copy the FP64 start-up result (V1, cross0) into box b of the fused state
int install_box(sh_ns_s* s, int b) {
    const long N = s->N;
    // the other boxes go back to (V1, N0) with the state's form
    RC(ns_canonical(s));
    s->w5 = false;
    for (int d = 0; d < 3; d++) {
        const long off = ((long)d * s->nboxes + b) * N;
        if (s->prec == SH_F64) {
            SH_HIP(hipMemcpyAsync((double2*)s->V1f + off, s->V1 + d * N, sizeof(double2) * N, hipMemcpyDeviceToDevice, s->st));
            SH_HIP(hipMemcpyAsync((double2*)s->N0f + off, s->cr0 + d * N, sizeof(double2) * N, hipMemcpyDeviceToDevice, s->st));
        } else {
            hipLaunchKernelGGL(k_z_to_c, dim3(egrid(N)), dim3(256), 0, s->st, s->V1 + d * N, (float2*)s->V1f + off, N);
            hipLaunchKernelGGL(k_z_to_c, dim3(egrid(N)), dim3(256), 0, s->st, s->cr0 + d * N, (float2*)s->N0f + off, N);
            SH_LAUNCH_CHECK("k_z_to_c");
        }
    }
    return SH_OK;
}

int ns_create(sh_ns_t* handle, int nx, int ny, int nz, int nboxes, const double* nus, double K,
              int precision, int fused, void* stream) {
    if (!handle || nx <= 0 || ny <= 0 || nz <= 0 || nboxes <= 0 || !nus ||
        (precision != SH_F64 && precision != SH_F32)) {
        set_error("sh_ns_create: invalid argument");
        return SH_EINVAL;
    }
    if (nboxes > 1 && !fused) {
        set_error("sh_ns_create_ensemble: ensembles use the fused schedule");
        return SH_EINVAL;
    }
    if (nboxes > 65535) {
        set_error("sh_ns_create_ensemble: too many boxes");
        return SH_EINVAL;
    }
    const double nu = nus[0];
    if (fused && !(fused_size_supported(nx) && fused_size_supported(ny) && fused_size_supported(nz))) {
        set_error("sh_ns_create: the fused schedule has no kernels for this grid (axis lengths 8 12 16 24 32 "
                  "48 64 96 128 256 512); the op-by-op schedule (fused = 0) runs any supported length");
        return SH_EUNSUPPORTED;
    }
    if (!fft_length_supported(nx) || !fft_length_supported(ny) || !fft_length_supported(nz)) {
        set_error("sh_ns_create: unsupported grid size");
        return SH_EUNSUPPORTED;
    }
    if (precision == SH_F32 && !fused) {
        set_error("sh_ns_create: the FP32 solver uses the fused schedule");
        return SH_EINVAL;
    }
    RC(ensure_init());
    sh_ns_s* s = new sh_ns_s();
    s->nx = nx;
    s->ny = ny;
    s->nz = nz;
    s->N = (long)nx * ny * nz;
    s->nu = nu;
    s->K = K;
    s->prec = precision;
    s->fused = fused;
    s->nonlinear = 1;
    s->st = as_stream(stream);
    s->device = current_device();
    s->have_div = false;
    s->w_primed = false;
    s->div_nb = 0;
    s->div_avg = 0;
    s->nboxes = nboxes;
    s->nus.assign(nus, nus + nboxes);
    s->coefs = nullptr;
    s->snap = nullptr;
    s->forcing = nullptr;
    s->cur_box = 0;
    s->dealias = 0;
    s->scheme = 0;
    s->layout = 1;
    s->half = false;
    s->box_herm.assign(nboxes, 1);
    s->overlap = 0;
    s->st2 = nullptr;
    s->ev_fork = s->ev_join = nullptr;
    s->cst = s->st;
    s->cx0 = 0;
    s->cnx = nx;
    const size_t z = sizeof(double2);
    int rc = 0;
    // FP64 state used by the start-up (and by the op-by-op loop)
    void* p;
#define AL(field, nf)                          \
    if (!rc) {                                 \
        rc = alloc_fields(s, &p, nf, z);       \
        s->field = (double2*)p;                \
    }
    AL(V0, 3) AL(V1, 3) AL(om0, 3) AL(om1, 3) AL(cr0, 3) AL(cr1, 3) AL(S, 3) AL(Gg, 3) AL(lap1, 3)
    AL(work, 6) AL(Pi0, 1)
    if (!fused) {
        AL(Vn, 3) AL(omn, 3) AL(Fv, 3) AL(Vh, 3) AL(G, 1) AL(div, 1)
    }
#undef AL
    if (!rc && fused) {
        const size_t e = precision == SH_F64 ? sizeof(double2) : sizeof(float2);
        rc = alloc_fields(s, &s->V1f, 3 * nboxes, e);
        if (!rc) rc = alloc_fields(s, &s->N0f, 3 * nboxes, e);
        if (!rc) rc = alloc_fields(s, &s->Wf, 6 * nboxes, e);
        s->divblocks = (int)((long)ny * nz + 1);  // >= blocks of the fused x kernel (>= 1 line per block)
        if (!rc) rc = pool_alloc((void**)&s->divpart, sizeof(double) * ((size_t)(s->divblocks + 1) * nboxes + 1));
        if (!rc) s->owned.push_back(s->divpart);
        if (!rc) rc = pool_alloc(&s->coefs, sizeof(StepCoef<double>) * nboxes);
        if (!rc) s->owned.push_back(s->coefs);
        if (!rc) rc = upload_coefs(s);
        // the overlapped schedule's second stream: non-blocking, so it never
        // waits on the legacy null stream implicitly; default priority
        if (!rc && hipStreamCreateWithFlags(&s->st2, hipStreamNonBlocking) != hipSuccess) rc = SH_EHIP;
        if (!rc && hipEventCreateWithFlags(&s->ev_fork, hipEventDisableTiming) != hipSuccess) rc = SH_EHIP;
        if (!rc && hipEventCreateWithFlags(&s->ev_join, hipEventDisableTiming) != hipSuccess) rc = SH_EHIP;
        if (rc == SH_EHIP) set_error("sh_ns_create: stream or event creation failed");
    }
    if (rc) {
        if (s->ev_fork) (void)hipEventDestroy(s->ev_fork);
        if (s->ev_join) (void)hipEventDestroy(s->ev_join);
        if (s->st2) (void)hipStreamDestroy(s->st2);
        for (void* q : s->owned) pool_free(q);
        delete s;
        return rc;
    }
    *handle = s;
    return SH_OK;
}

}  // namespace

extern "C" {

int sh_ns_create(sh_ns_t* handle, int nx, int ny, int nz, double nu, double K, int precision,
                 int fused, void* stream) {
    return ns_create(handle, nx, ny, nz, 1, &nu, K, precision, fused, stream);
}

int sh_ns_create_ensemble(sh_ns_t* handle, int n, int nboxes, const double* nu, double K,
                          int precision, void* stream) {
    return ns_create(handle, n, n, n, nboxes, nu, K, precision, 1, stream);
}

int sh_ns_box_count(sh_ns_t s, int* count) {
    if (!s || !count) return SH_EINVAL;
    *count = s->nboxes;
    return SH_OK;
}

// V0 of box b (Fourier, 3 contiguous complex128 fields) followed by that box's
// start-up (exec/ns.cpp:132-153) with its own viscosity
int sh_ns_set_initial_box(sh_ns_t s, int box, const void* V0) {
    if (!s || !V0 || box < 0 || box >= s->nboxes) return SH_EINVAL;
    // the other boxes go back to (V1, N0) while W still holds their x-inverse
    RC(ns_canonical(s));
    SH_HIP(hipMemcpyAsync(s->V0, V0, sizeof(double2) * 3 * s->N, hipMemcpyDeviceToDevice, s->st));
    RC(curl(s, s->V0, s->om0));
    const double nu0 = s->nu;
    s->nu = s->nus[box];
    s->cur_box = box;
    int rc = startup64(s);
    s->nu = nu0;
    s->cur_box = 0;
    if (rc) return rc;
    RC(decide_layout(s, box));
    s->w_primed = false;
    return install_box(s, box);
}

int sh_ns_get_velocity_box(sh_ns_t s, int box, void* out) {
    if (!s || !out || box < 0 || box >= s->nboxes || !s->fused) return SH_EINVAL;
    const long N = s->N;
    if (s->yform) {  // V1 of the box from W
        RC(v1_into_z(s, (double2*)out, box));
        return s->half ? expand_half<double>(s, out, 3) : SH_OK;
    }
    for (int d = 0; d < 3; d++) {
        const long off = ((long)d * s->nboxes + box) * N;
        if (s->prec == SH_F64) {
            SH_HIP(hipMemcpyAsync((double2*)out + d * N, (double2*)s->V1f + off, sizeof(double2) * N,
                                  hipMemcpyDeviceToDevice, s->st));
        } else {
            hipLaunchKernelGGL(k_c_to_z, dim3(egrid(N)), dim3(256), 0, s->st, (const float2*)s->V1f + off,
                               (double2*)out + d * N, N);
            SH_LAUNCH_CHECK("k_c_to_z");
        }
    }
    if (s->half) return expand_half<double>(s, out, 3);
    return SH_OK;
}

}  // extern "C"

namespace {
template <typename T>
int snapshot_t(sh_ns_s* s, int r, float* out) {
    const long N = s->N;
    const int cx = s->nx / r, cy = s->ny / r, cz = s->nz / r;
    const long nc = (long)cx * cy * cz;
    // all boxes at once: one copy, one 3-pass inverse transform, one filter launch
    C2<T>* snap = (C2<T>*)s->snap;
    const long nf = 3L * s->nboxes;
    if (s->yform)  // V1 of every box from W (the state stays as it is)
        RC(v1_into<T>(s, snap));
    else
        SH_HIP(hipMemcpyAsync(snap, s->V1f, sizeof(C2<T>) * N * nf, hipMemcpyDeviceToDevice, s->st));
    if (s->half) RC(expand_half<T>(s, snap, (int)nf));
    RC(fft3d<T>(snap, (int)nf, N, s->nx, s->ny, s->nz, +1, SH_NORMALIZE, s->st));
    unsigned gx = egrid(nc);
    hipLaunchKernelGGL(k_box_filter<T>, dim3(gx, (unsigned)s->nboxes), dim3(256), 0, s->st, snap, N,
                       N * s->nboxes, s->nx, s->ny, r, cx, cy, nc, out);
    SH_LAUNCH_CHECK("k_box_filter");
    return SH_OK;
}

// snapshot scratch: 6 fields per box (velocity and vorticity), first use only
int ensure_snap(sh_ns_s* s) {
    if (s->snap) return SH_OK;
    const size_t e = s->prec == SH_F64 ? sizeof(double2) : sizeof(float2);
    RC(pool_alloc(&s->snap, 6 * e * (size_t)s->N * s->nboxes));
    s->owned.push_back(s->snap);
    return SH_OK;
}

template <typename T>
int snapshot_force_t(sh_ns_s* s, int r, float* out) {
    const long N = s->N;
    const int cx = s->nx / r, cy = s->ny / r, cz = s->nz / r;
    const long nc = (long)cx * cy * cz;
    const long nf = 3L * s->nboxes;
    C2<T>* u = (C2<T>*)s->snap;
    C2<T>* w = u + nf * N;
    if (s->yform)
        RC(v1_into<T>(s, u));
    else
        SH_HIP(hipMemcpyAsync(u, s->V1f, sizeof(C2<T>) * N * nf, hipMemcpyDeviceToDevice, s->st));
    if (s->half) RC(expand_half<T>(s, u, (int)nf));
    hipLaunchKernelGGL(k_curl_boxes<T>, dim3(egrid(N), (unsigned)s->nboxes), dim3(256), 0, s->st, u, w, N,
                       N * s->nboxes, s->nx, s->ny, s->nz);
    SH_LAUNCH_CHECK("k_curl_boxes");
    RC(fft3d<T>(u, (int)(2 * nf), N, s->nx, s->ny, s->nz, +1, SH_NORMALIZE, s->st));
    hipLaunchKernelGGL(k_box_cross<T>, dim3(egrid(nc), (unsigned)s->nboxes), dim3(256), 0, s->st, u, w, N,
                       N * s->nboxes, s->nx, s->ny, r, cx, cy, nc, out);
    SH_LAUNCH_CHECK("k_box_cross");
    return SH_OK;
}
}  // namespace

extern "C" {

// Box-filtered snapshot of every box (training data): physical velocity by
// an inverse transform, then ratio^3 averages.  out: float[nboxes][9][nc],
// nc = (nx/ratio)(ny/ratio)(nz/ratio); fields ubar_x ubar_y ubar_z and
// tau_xx tau_yy tau_zz tau_xy tau_xz tau_yz.  The first call allocates a
// scratch buffer, so make it before any graph capture.
int sh_ns_snapshot(sh_ns_t s, int ratio, void* out) {
    if (!s || !out || !s->fused || ratio <= 0 || s->nx % ratio || s->ny % ratio || s->nz % ratio) {
        set_error("sh_ns_snapshot: invalid argument (fused solver, ratio dividing the grid)");
        return SH_EINVAL;
    }
    RC(ensure_snap(s));
    return s->prec == SH_F64 ? snapshot_t<double>(s, ratio, (float*)out) : snapshot_t<float>(s, ratio, (float*)out);
}

// Box-filtered nonlinear term of every box: the ratio^3 averages of the
// physical u x omega.  out: float[nboxes][3][nc].  Shares the snapshot scratch.
int sh_ns_snapshot_force(sh_ns_t s, int ratio, void* out) {
    if (!s || !out || !s->fused || ratio <= 0 || s->nx % ratio || s->ny % ratio || s->nz % ratio) {
        set_error("sh_ns_snapshot_force: invalid argument (fused solver, ratio dividing the grid)");
        return SH_EINVAL;
    }
    RC(ensure_snap(s));
    return s->prec == SH_F64 ? snapshot_force_t<double>(s, ratio, (float*)out)
                             : snapshot_force_t<float>(s, ratio, (float*)out);
}

int sh_ns_div_average_box(sh_ns_t s, int box, double* result) {
    if (!s || !result || box < 0 || box >= s->nboxes || !s->fused) return SH_EINVAL;
    if (!s->have_div) {
        *result = 0;
        return SH_OK;
    }
    double* part = s->divpart + (long)box * (s->divblocks + 1);
    hipLaunchKernelGGL(k_abs_sum_partial, dim3(1), dim3(256), 0, s->st, part, s->div_nb, part + s->divblocks);
    SH_LAUNCH_CHECK("k_abs_sum_partial");
    double v = 0;
    SH_HIP(hipMemcpyAsync(&v, part + s->divblocks, sizeof(double), hipMemcpyDeviceToHost, s->st));
    SH_HIP(hipStreamSynchronize(s->st));
    *result = v / (double)s->N;
    return SH_OK;
}

int sh_ns_destroy(sh_ns_t s) {
    if (!s) return SH_OK;
    (void)hipStreamSynchronize(s->st);
    if (s->st2) {
        (void)hipStreamSynchronize(s->st2);
        (void)hipStreamDestroy(s->st2);
    }
    if (s->ev_fork) (void)hipEventDestroy(s->ev_fork);
    if (s->ev_join) (void)hipEventDestroy(s->ev_join);
    for (void* q : s->owned) pool_free(q);
    delete s;
    return SH_OK;
}

int sh_ns_set_forcing(sh_ns_t s, const void* f3) {
    if (!s) return SH_EINVAL;
    s->forcing = f3;
    return SH_OK;
}

int sh_ns_set_scheme(sh_ns_t s, int scheme) {
    if (!s || scheme < 0 || scheme > 1) {
        set_error("sh_ns_set_scheme: scheme must be 0 (exec/ns.cpp) or 1 (exec/ns2.cpp)");
        return SH_EINVAL;
    }
    if (s->fused) RC(ns_canonical(s));  // Y holds the old scheme's viscous factor
    s->scheme = scheme;
    return s->fused ? upload_coefs(s) : SH_OK;
}

int sh_ns_set_dealias(sh_ns_t s, int mode) {
    if (!s || mode < 0 || mode > 2) {
        set_error("sh_ns_set_dealias: mode must be 0, 1 or 2");
        return SH_EINVAL;
    }
    s->dealias = mode;
    return s->fused ? upload_coefs(s) : SH_OK;
}

int sh_ns_set_nonlinear(sh_ns_t s, int on) {
    if (!s) return SH_EINVAL;
    s->nonlinear = on != 0;
    return SH_OK;
}

int sh_ns_set_initial(sh_ns_t s, const void* V0) {
    if (!s || !V0) return SH_EINVAL;
    SH_HIP(hipMemcpyAsync(s->V0, V0, sizeof(double2) * 3 * s->N, hipMemcpyDeviceToDevice, s->st));
    return curl(s, s->V0, s->om0);
}

int sh_ns_set_vorticity(sh_ns_t s, const void* omega0) {
    if (!s || !omega0) return SH_EINVAL;
    SH_HIP(hipMemcpyAsync(s->om0, omega0, sizeof(double2) * 3 * s->N, hipMemcpyDeviceToDevice, s->st));
    return SH_OK;
}

int sh_ns_startup(sh_ns_t s) {
    if (!s) return SH_EINVAL;
    if (s->nboxes > 1) {
        set_error("sh_ns_startup: an ensemble starts each box in sh_ns_set_initial_box");
        return SH_EINVAL;
    }
    s->w_primed = false;
    RC(startup64(s));
    if (!s->fused) return SH_OK;
    RC(decide_layout(s, 0));
    return install_box(s, 0);
}

int sh_ns_steps(sh_ns_t s, int nsteps) {
    if (!s || nsteps < 0) return SH_EINVAL;
    for (int t = 0; t < nsteps; t++) {
        if (!s->fused)
            RC(step64(s));
        else if (s->prec == SH_F64)
            RC(step_fused<double>(s));
        else
            RC(step_fused<float>(s));
    }
    return SH_OK;
}

// the single-domain getters; an ensemble's boxes are field-major (field d of
// box b at (d*nboxes + b)*N), so they are read through the *_box calls
static int reject_ensemble(sh_ns_t s, const char* who) {
    if (s->nboxes == 1) return SH_OK;
    set_error(std::string(who) + ": the handle holds " + std::to_string(s->nboxes) +
              " boxes; use the *_box call with a box index");
    return SH_EINVAL;
}

int sh_ns_get_velocity(sh_ns_t s, void* out) {
    if (!s || !out) return SH_EINVAL;
    RC(reject_ensemble(s, "sh_ns_get_velocity"));
    const long n3 = 3 * s->N;
    if (!s->fused) {
        SH_HIP(hipMemcpyAsync(out, s->V1, sizeof(double2) * n3, hipMemcpyDeviceToDevice, s->st));
    } else if (s->yform) {
        RC(v1_into_z(s, (double2*)out));
    } else if (s->prec == SH_F64) {
        SH_HIP(hipMemcpyAsync(out, s->V1f, sizeof(double2) * n3, hipMemcpyDeviceToDevice, s->st));
    } else {
        hipLaunchKernelGGL(k_c_to_z, dim3(egrid(n3)), dim3(256), 0, s->st, (const float2*)s->V1f, (double2*)out, n3);
        SH_LAUNCH_CHECK("k_c_to_z");
    }
    if (s->fused && s->half) return expand_half<double>(s, out, 3);
    return SH_OK;
}

int sh_ns_set_layout(sh_ns_t s, int mode) {
    if (!s || mode < 0 || mode > 1) {
        set_error("sh_ns_set_layout: mode must be 0 (full spectrum) or 1 (half when Hermitian)");
        return SH_EINVAL;
    }
    s->layout = mode;
    if (mode == 0 && s->half) {  // leave the half layout at once
        RC(expand_state(s));
        s->half = false;
        s->w_primed = false;
    }
    return SH_OK;
}

int sh_ns_set_overlap(sh_ns_t s, int on) {
    if (!s || on < 0 || on > 1) {
        set_error("sh_ns_set_overlap: 0 or 1");
        return SH_EINVAL;
    }
    if (on && !s->fused) {
        set_error("sh_ns_set_overlap: only the fused schedule has an overlapped form");
        return SH_EINVAL;
    }
    s->overlap = on;
    return SH_OK;
}

int sh_ns_get_layout(sh_ns_t s, int* half) {
    if (!s || !half) return SH_EINVAL;
    *half = s->half ? 1 : 0;
    return SH_OK;
}

int sh_ns_get_history_form(sh_ns_t s, int* form) {
    if (!s || !form) return SH_EINVAL;
    *form = s->yform ? (s->w5 ? 2 : 1) : 0;
    return SH_OK;
}

int sh_ns_prepare(sh_ns_t s) {
    if (!s) return SH_EINVAL;
    if (!s->fused) return SH_OK;
    if (capturing(s->st)) {
        set_error("sh_ns_prepare: not inside a graph capture");
        return SH_EINVAL;
    }
    const bool f64 = s->prec == SH_F64;
    const bool merged = fourier_merged(s);
    if (s->yform && !merged) RC(ns_canonical(s));
    if (!s->w_primed) {  // kernel A, as a step's first kernel would run it
        RC(f64 ? dispatch_inv_x<double>(s) : dispatch_inv_x<float>(s));
        s->w_primed = true;
    }
    if (!s->yform && merged) {
        const int rc = f64 ? launch_yform<double, true>(s) : launch_yform<float, true>(s);
        if (rc) return rc;
        s->yform = true;
    }
    return SH_OK;
}

int sh_ns_get_vorticity(sh_ns_t s, void* out) {
    if (!s || !out) return SH_EINVAL;
    RC(reject_ensemble(s, "sh_ns_get_vorticity"));
    if (!s->fused) {
        SH_HIP(hipMemcpyAsync(out, s->om1, sizeof(double2) * 3 * s->N, hipMemcpyDeviceToDevice, s->st));
        return SH_OK;
    }
    // omega1 = curl(V1) (exec/ns.cpp:206, 216)
    RC(sh_ns_get_velocity(s, s->work));
    return curl(s, s->work, (double2*)out);
}

int sh_ns_div_average(sh_ns_t s, double* result) {
    if (!s || !result) return SH_EINVAL;
    RC(reject_ensemble(s, "sh_ns_div_average"));
    if (!s->have_div) {
        *result = 0;
        return SH_OK;
    }
    if (!s->fused) return sh_average_z(s->div, s->N, result, s->st);
    // the last kernel E wrote one partial per block (div_nb <= divblocks)
    hipLaunchKernelGGL(k_abs_sum_partial, dim3(1), dim3(256), 0, s->st, s->divpart, s->div_nb,
                       s->divpart + s->divblocks);
    SH_LAUNCH_CHECK("k_abs_sum_partial");
    double v = 0;
    SH_HIP(hipMemcpyAsync(&v, s->divpart + s->divblocks, sizeof(double), hipMemcpyDeviceToHost, s->st));
    SH_HIP(hipStreamSynchronize(s->st));
    *result = v / (double)s->N;
    return SH_OK;
}

int sh_ns_state_bytes(sh_ns_t s, size_t* bytes) {
    if (!s || !bytes) return SH_EINVAL;
    // every owned block was allocated through the pool with its exact size request
    size_t fields = 0;
    fields += 3 * 6 + 9 + 6 + 1;  // FP64 start-up state
    if (!s->fused) fields += 3 * 4 + 2;
    size_t b = fields * sizeof(double2) * (size_t)s->N;
    if (s->fused) b += 12 * (s->prec == SH_F64 ? sizeof(double2) : sizeof(float2)) * (size_t)s->N;
    *bytes = b;
    return SH_OK;
}

}  // extern "C"

namespace sh {
// the fused solver's current velocity (3 Fourier fields per box, precision prec)
void* ns_fused_velocity(sh_ns_t s) {
    if (!s || !s->fused || ns_canonical(s) != SH_OK) return nullptr;
    return s->V1f;
}

// V1 is read in place by the caller from now on: the state keeps (V1, N0)
void ns_keep_velocity(sh_ns_t s) {
    if (!s) return;
    s->keep_v1 = true;
    (void)ns_canonical(s);
}
}  // namespace sh

namespace sh {
// one fused step in two parts: phase 0 = kernels A-D, phase 1 = kernel E
// (the only one that reads the body force), so a caller can compute the
// force concurrently with phase 0
int ns_step_phase(sh_ns_t s, int phase) {
    if (!s || !s->fused) return SH_EINVAL;
    if (phase == 0) return s->prec == SH_F64 ? step_fused_front<double>(s) : step_fused_front<float>(s);
    int rc = s->prec == SH_F64 ? dispatch_fourier_x<double>(s) : dispatch_fourier_x<float>(s);
    if (!rc) s->have_div = true;
    return rc;
}
}  // namespace sh
