// les.hip -- coarse-grid LES with the learned subgrid-stress closure.  No
// counterpart in the reference.
//
// Each step evaluates the closure on the current velocity V (Fourier, FP32)
// and hands f = -div tau to the fused solver as its body force:
//   1  g_{3i+j} = i k_j V_i, packed two real fields per complex field
//      (Z_m = g_{2m} + i g_{2m+1}, m = 0..4, Z_4 = g_8): 5 fields, not 9
//   2  inverse 3D transform of Z (5 fields)
//   3  BF16 MFMA network on the packed physical gradients -> packed stresses
//      (tau_xx + i tau_yy, tau_zz + i tau_xy, tau_xz + i tau_yz)
//   4  forward z and y passes of the stresses (3 fields, not 6); the z passes
//      of steps 2 and 4 and the network run as one kernel (k_les_z_mlp)
//   5  the x pass on pairs of mirror lines, fused with the unpacking by
//      Hermitian symmetry and f_i = -sum_j i k_j tau_ij
// Nyquist wavenumbers count as zero in the derivatives, so every packed field
// is the transform of real fields and the packing is exact.
// With the lagged closure (sh_les_set_lag) step t+1 is forced by an
// extrapolation from the closures of steps t and t-1, and the closure chain
// runs beside the solver's whole step instead of its kernels A-D.
#include <cmath>

#include "fft_core.hpp"
#include "mlp_core.hpp"

namespace sh {

template <typename T>
int fft3d(void* data, int count, long fstride, int nx, int ny, int nz, int dir, int flags, void* stream);
int mlp_infer_packed(sh_mlp_t h, const float2* X, long ld, long n, float2* Y, long ldy, int clip, float smag,
                     hipStream_t st);
int mlp_infer_packed_planar(sh_mlp_t h, const float2* X, long ld, long n, float* Y, long ldy, int clip, float smag,
                            hipStream_t st);
void* ns_fused_velocity(sh_ns_t s);
void ns_keep_velocity(sh_ns_t s);
int ns_step_phase(sh_ns_t s, int phase);
int ns_autotune_with(sh_ns_t s, int reps, int (*step_cb)(void*), void* ctx);
int mlp_infer_planar_packed(sh_mlp_t h, const float* X, long ld, long n, float2* Y, long ldy, hipStream_t st);
int mlp_grad_dy(sh_mlp_t h, const float* X, const float* Dy, long ld, long n, hipStream_t st);
int mlp_adam(sh_mlp_t h, float lr, hipStream_t st);
void* mlp_aux(sh_mlp_t h, size_t bytes);
float* mlp_grad_buffer(sh_mlp_t h);
const float* mlp_scales(sh_mlp_t h);
const float* mlp_params(sh_mlp_t h);
hipStream_t mlp_stream(sh_mlp_t h);

// wavenumber of index i on an n-point axis, with the Nyquist index as zero
__device__ __forceinline__ float kd(int i, int n) { return (2 * i == n) ? 0.f : (float)(i <= n / 2 ? i : i - n); }

__device__ __forceinline__ float2 ik(float k, float2 v) { return make_float2(-k * v.y, k * v.x); }

// exp(-36 (|k| / (n/2))^36), the solver's exponential dealiasing factor of one axis
__device__ __forceinline__ float expfilter_f(float k, int n) {
    const double r = fabs((double)k) / (0.5 * n);
    const double r4 = (r * r) * (r * r);
    const double r8 = r4 * r4, r16 = r8 * r8;
    return (float)exp(-36.0 * (r16 * r16) * r4);  // r^32 r^4 = r^36
}

// Step 1 (Z_m = g_{2m} + i g_{2m+1}, g_{3a+b} = i k_b V_a) fused with the x
// pass of step 2: block (x, m) forms packed field m of LPB x-lines on load
// (the derivative factors of line (j, k) and of each position) and
// transforms it back along x in registers and LDS.  half = 1: V holds only
// the planes k < n/2 (the solver's half-spectrum layout of a real field);
// a line with k > n/2 is read from its mirror line (-j, -k) as
// V(kx, ky, kz) = conj V(-kx, -ky, -kz), and the plane k = n/2 is zero.
template <class P, int LPB>
__global__ __launch_bounds__(P::TPL* LPB) void k_les_grad_x(const float2* __restrict__ V, float2* __restrict__ Z,
                                                            const float2* __restrict__ tw, int n, int half) {
    constexpr int N = P::N;
    constexpr int LS = N + N / 16 + 1;
    __shared__ float sre[LPB * LS];
    __shared__ float sim[LPB * LS];
    const int tid = threadIdx.x % P::TPL;
    const int l = threadIdx.x / P::TPL;
    const long nlines = (long)n * n;
    const long N3 = nlines * N;
    const long line = (long)blockIdx.x * LPB + l;
    const bool valid = line < nlines;
    const long lc = valid ? line : 0;
    const int m = blockIdx.y;
    // components 2m and 2m+1 of g_{3a+b} = i k_b V_a (the second is absent for m = 4)
    const int a0 = (2 * m) / 3, b0 = (2 * m) % 3, a1 = (2 * m + 1) / 3, b1 = (2 * m + 1) % 3;
    const bool two = m < 4;
    const int jl = (int)(lc % n), kl = (int)(lc / n);
    const float ky = kd(jl, n), kz = kd(kl, n);
    // where the line's values are read: itself, or (half layout, k > n/2)
    // its mirror line, conjugated with kx reversed; nothing on the plane n/2
    const bool mirror = half && 2 * kl > n;
    const bool zero = half && 2 * kl == n;
    const long src = mirror ? (long)(n - kl) * n + (n - jl) % n : lc;
    const float2* v0 = V + a0 * N3 + src * N;
    const float2* v1 = V + (two ? a1 : a0) * N3 + src * N;
    auto vget = [&](const float2* v, int pos) {
        if (zero) return make_float2(0.f, 0.f);
        if (!mirror) return v[pos];
        const float2 w = v[(N - pos) % N];
        return make_float2(w.x, -w.y);
    };
    float2 x[P::E];
    for_in<P>(tid, [&](int s, int pos) {
        const float kx = kd(pos, N);
        const float kb0 = b0 == 0 ? kx : b0 == 1 ? ky : kz;
        const float2 g0 = ik(kb0, vget(v0, pos));
        float2 z = g0;
        if (two) {
            const float kb1 = b1 == 0 ? kx : b1 == 1 ? ky : kz;
            const float2 g1 = ik(kb1, vget(v1, pos));
            z = make_float2(g0.x - g1.y, g0.y + g1.x);
        }
        x[s] = z;
    });
    auto idx = [&](int pos) { return l * LS + pos + (pos >> 4); };
    line_fft<P, 1, float>(x, tid, sre, sim, idx, tw);
    if (!valid) return;
    float2* zb = Z + m * N3 + lc * N;
    for_out<P>(tid, [&](int s, int pos) { zb[pos] = x[s]; });
}

// Steps 1 and 2's x and y passes in one kernel for one z-plane k of one packed
// field m (block (k, m)): Z_m formed on load as k_les_grad_x does (with its
// half-layout mirror reads), the inverse x transforms of the plane's n rows,
// then the inverse y transforms of its n columns, through the plane held in
// LDS (element (y, x) at y*LS + x + x/16); the plane is stored row by row.
// Rows run with a line's threads adjacent, columns with adjacent columns, so
// both phases address LDS without bank conflicts.
template <class P>
__global__ __launch_bounds__(P::TPL* P::N) void k_les_grad_xy(const float2* __restrict__ V, float2* __restrict__ Z,
                                                              const float2* __restrict__ tw, int n, int half) {
    constexpr int N = P::N;
    constexpr int LS = N + N / 16 + 1;
    __shared__ float pre[N * LS];
    __shared__ float pim[N * LS];
    const long N3 = (long)N * N * N;
    const int kl = blockIdx.x, m = blockIdx.y;
    auto at = [](int y, int x) { return y * LS + x + (x >> 4); };
    float2 x[P::E];
    {  // the x transform of row j = line (j, kl)
        const int tid = threadIdx.x % P::TPL, jl = threadIdx.x / P::TPL;
        const int a0 = (2 * m) / 3, b0 = (2 * m) % 3, a1 = (2 * m + 1) / 3, b1 = (2 * m + 1) % 3;
        const bool two = m < 4;
        const float ky = kd(jl, n), kz = kd(kl, n);
        const bool mirror = half && 2 * kl > n;
        const bool zero = half && 2 * kl == n;
        const long src = mirror ? (long)(n - kl) * n + (n - jl) % n : (long)kl * n + jl;
        const float2* v0 = V + a0 * N3 + src * N;
        const float2* v1 = V + (two ? a1 : a0) * N3 + src * N;
        auto vget = [&](const float2* v, int pos) {
            if (zero) return make_float2(0.f, 0.f);
            if (!mirror) return v[pos];
            const float2 w = v[(N - pos) % N];
            return make_float2(w.x, -w.y);
        };
        for_in<P>(tid, [&](int s, int pos) {
            const float kx = kd(pos, N);
            const float kb0 = b0 == 0 ? kx : b0 == 1 ? ky : kz;
            const float2 g0 = ik(kb0, vget(v0, pos));
            float2 z = g0;
            if (two) {
                const float kb1 = b1 == 0 ? kx : b1 == 1 ? ky : kz;
                const float2 g1 = ik(kb1, vget(v1, pos));
                z = make_float2(g0.x - g1.y, g0.y + g1.x);
            }
            x[s] = z;
        });
        auto ridx = [&](int pos) { return at(jl, pos); };
        line_fft<P, 1, float>(x, tid, pre, pim, ridx, tw);
        __syncthreads();
        for_out<P>(tid, [&](int s, int pos) {
            pre[ridx(pos)] = x[s].x;
            pim[ridx(pos)] = x[s].y;
        });
    }
    __syncthreads();
    {  // the y transform of column i, in place in the plane
        const int col = threadIdx.x % N, tid = threadIdx.x / N;
        auto cidx = [&](int pos) { return at(pos, col); };
        for_in<P>(tid, [&](int s, int pos) { x[s] = make_float2(pre[cidx(pos)], pim[cidx(pos)]); });
        __syncthreads();  // a column's threads are in different waves: all read before any writes
        line_fft<P, 1, float>(x, tid, pre, pim, cidx, tw);
        __syncthreads();
        for_out<P>(tid, [&](int s, int pos) {
            pre[cidx(pos)] = x[s].x;
            pim[cidx(pos)] = x[s].y;
        });
    }
    __syncthreads();
    float2* zp = Z + m * N3 + (long)kl * N * N;
    for (int e = threadIdx.x; e < N * N; e += blockDim.x) {
        const int y = e / N, xx = e % N;
        zp[e] = make_float2(pre[at(y, xx)], pim[at(y, xx)]);
    }
}

// f_a = -sum_b i k_b tau_ab from the packed stress transforms P_c:
// tau_{2c} = (P_c(k) + conj P_c(-k)) / 2, tau_{2c+1} = (P_c(k) - conj P_c(-k)) / 2i
__global__ void k_les_div_unpack(const float2* __restrict__ Pt, long ts, float2* __restrict__ F, long fs, int n) {
    const long N = (long)n * n * n;
    const int pair[3][3] = {{0, 3, 4}, {3, 1, 5}, {4, 5, 2}};  // xx yy zz xy xz yz
    for (long p = blockIdx.x * (long)blockDim.x + threadIdx.x; p < N; p += (long)gridDim.x * blockDim.x) {
        const int i = (int)(p % n), j = (int)((p / n) % n), k = (int)(p / ((long)n * n));
        const long m = (long)((n - i) % n) + (long)n * ((n - j) % n) + (long)n * n * ((n - k) % n);
        float2 t[6];
        for (int c = 0; c < 3; c++) {
            const float2 a = Pt[c * ts + p], b = Pt[c * ts + m];
            t[2 * c] = make_float2(0.5f * (a.x + b.x), 0.5f * (a.y - b.y));
            t[2 * c + 1] = make_float2(0.5f * (a.y + b.y), -0.5f * (a.x - b.x));
        }
        const bool nyq = 2 * i == n || 2 * j == n || 2 * k == n;
        const float kk[3] = {kd(i, n), kd(j, n), kd(k, n)};
        for (int a = 0; a < 3; a++) {
            float2 s = make_float2(0.f, 0.f);
            for (int b = 0; b < 3; b++) {
                const float2 d = ik(kk[b], t[pair[a][b]]);
                s.x += d.x;
                s.y += d.y;
            }
            F[a * fs + p] = nyq ? make_float2(0.f, 0.f) : make_float2(-s.x, -s.y);
        }
    }
}

// x-line pair q of an n x n grid of (y, z) lines: the lines (j, k) and
// (-j, -k) hold the modes k and -k.  Pairs 0 .. (n/2-1) n - 1 are the rows
// k = 1 .. n/2-1 (every j), the last 2 (n/2 + 1) the rows k = 0 and k = n/2
// with j <= n/2 (those rows are their own mirror).  n even.
__device__ __forceinline__ void pair_line(long q, int n, int& j, int& k) {
    const long mid = (long)(n / 2 - 1) * n;
    if (q < mid) {
        k = 1 + (int)(q / n);
        j = (int)(q % n);
        return;
    }
    const int r = (int)(q - mid);
    k = r <= n / 2 ? 0 : n / 2;
    j = r <= n / 2 ? r : r - (n / 2 + 1);
}

// The last pass of the stress transform fused with the unpacking and the
// divergence: the forward x pass of the packed stresses P_c (c = 0..2, y and
// z already forward-transformed) on a line and its mirror line, then
//   tau_{2c} = (P_c(k) + conj P_c(-k)) / 2, tau_{2c+1} = (P_c(k) - conj P_c(-k)) / 2i,
//   f_a = -sum_b i k_b tau_ab   (zero on the Nyquist planes).
// mode 0: F = f;  1: F = H = f;  2 (lagged closure): F = 2 f - H, H = f, so F
// extrapolates the force of the next step from this one and the previous one.
template <class P, int PPB>
__global__ __launch_bounds__(P::TPL * 6 * PPB) void k_les_x_div(const float2* __restrict__ T,
                                                                float2* __restrict__ F,
                                                                float2* __restrict__ H,
                                                                const float2* __restrict__ tw, int n, int mode) {
    constexpr int N = P::N;
    constexpr int LS = N + N / 16 + 1;
    __shared__ float sre[6 * PPB * LS];
    __shared__ float sim[6 * PPB * LS];
    const int tid = threadIdx.x % P::TPL;
    const int ll = threadIdx.x / P::TPL;  // 0 .. 6 PPB - 1: pair ll / 6, group ll % 6
    const int g = ll % 6;                 // 0..2: field g of the line, 3..5: field g - 3 of its mirror
    const long N3 = (long)n * n * n;
    const long npairs = (long)n * n / 2 + 2;
    const long q0 = (long)blockIdx.x * PPB + ll / 6;
    int j, k;
    pair_line(q0 < npairs ? q0 : 0, n, j, k);
    const long line = g < 3 ? (long)k * n + j : (long)((n - k) % n) * n + (n - j) % n;
    const float2* src = T + (g % 3) * N3 + line * N;
    float2 x[P::E];
    for_in<P>(tid, [&](int s, int pos) { x[s] = src[pos]; });
    auto idx = [&](int grp, int pos) { return grp * LS + pos + (pos >> 4); };
    auto lidx = [&](int pos) { return idx(ll, pos); };
    line_fft<P, -1, float>(x, tid, sre, sim, lidx, tw);
    __syncthreads();
    for_out<P>(tid, [&](int s, int pos) {
        sre[lidx(pos)] = x[s].x;
        sim[lidx(pos)] = x[s].y;
    });
    __syncthreads();
    const int pair[3][3] = {{0, 3, 4}, {3, 1, 5}, {4, 5, 2}};  // xx yy zz xy xz yz
    for (int o = threadIdx.x; o < PPB * 2 * N; o += blockDim.x) {
        const int p2 = o / (2 * N), side = (o / N) % 2, i = o % N;
        const long q = (long)blockIdx.x * PPB + p2;
        if (q >= npairs) continue;
        int jj, kk;
        pair_line(q, n, jj, kk);
        const int jm = (n - jj) % n, km = (n - kk) % n;
        if (side == 1 && jm == jj && km == kk) continue;  // a line that is its own mirror
        const int lj = side ? jm : jj, lk = side ? km : kk;
        const int im = (N - i) % N;
        float2 t[6];
        for (int c = 0; c < 3; c++) {
            const int ga = 6 * p2 + 3 * side + c, gb = 6 * p2 + 3 * (1 - side) + c;
            const float2 a = make_float2(sre[idx(ga, i)], sim[idx(ga, i)]);
            const float2 b = make_float2(sre[idx(gb, im)], sim[idx(gb, im)]);
            t[2 * c] = make_float2(0.5f * (a.x + b.x), 0.5f * (a.y - b.y));
            t[2 * c + 1] = make_float2(0.5f * (a.y + b.y), -0.5f * (a.x - b.x));
        }
        const bool nyq = 2 * i == n || 2 * lj == n || 2 * lk == n;
        const float kk3[3] = {kd(i, n), kd(lj, n), kd(lk, n)};
        const long p = ((long)lk * n + lj) * N + i;
        for (int a = 0; a < 3; a++) {
            float2 s = make_float2(0.f, 0.f);
            for (int b = 0; b < 3; b++) {
                const float2 d = ik(kk3[b], t[pair[a][b]]);
                s.x += d.x;
                s.y += d.y;
            }
            const float2 f = nyq ? make_float2(0.f, 0.f) : make_float2(-s.x, -s.y);
            if (mode == 2) {
                const float2 h = H[a * N3 + p];
                F[a * N3 + p] = make_float2(2.f * f.x - h.x, 2.f * f.y - h.y);
            } else {
                F[a * N3 + p] = f;
            }
            if (mode >= 1) H[a * N3 + p] = f;
        }
    }
}

// The closure's z passes and the network in one kernel, for LPB x-columns of
// one y-row (every z), with 8 thread groups of one wave each: the inverse z
// transforms of the 5 packed gradient fields (groups 0-4, scaled by 1/n^3),
// the network on the block's LPB*n grid points with all 8 waves (staged in
// LDS as [feature][sample], sample pos*LPB + column), and the forward z
// transforms of the 3 packed stress fields (groups 0-2).  The other groups
// transform zeros in scratch of their own, so that every thread meets the
// same barriers.  The BF16 weight operands are formed once per block in LDS.
// The separate launches with the default z plan (P below) and k_mlp_infer
// do the same operations; where the compiler rounds the transforms the same
// way in both (measured at n = 16 and 32) the results are the same bit for
// bit, otherwise they agree to FP32 rounding (a few BF16 activations round
// the other way).
// forward2 (mlp_core.hpp) with the weight operands and biases read from LDS
// at each use: operand o of lane l at W[64 o + l] (W1^T tile mt at o = mt,
// W2^T tile (mt, kb) at 4 + 4 mt + kb, W3^T tile kb at 20 + kb), bias tile k
// of lane group q at B[4 k + q] (b1: k = mt, b2: 4 + mt, b3: 8).  The same
// operations in the same order, so the same results.
__device__ __forceinline__ void forward2_lds(const s16x4* W, const f32x4* B, int lane, int q, s16x4 xa, s16x4 xb,
                                             f32x4& ya, f32x4& yb) {
    s16x4 ha[4], hb[4];
    for (int mt = 0; mt < 4; mt++) {
        const s16x4 w = W[mt * 64 + lane];
        const f32x4 b = B[4 * mt + q];
        ha[mt] = to_bf(tanh4(mfma(w, xa, b)));
        hb[mt] = to_bf(tanh4(mfma(w, xb, b)));
    }
    s16x4 ga[4], gb[4];
    for (int mt = 0; mt < 4; mt++) {
        f32x4 a = B[4 * (4 + mt) + q], c = a;
        for (int kb = 0; kb < 4; kb++) {
            const s16x4 w = W[(4 + 4 * mt + kb) * 64 + lane];
            a = mfma(w, ha[kb], a);
            c = mfma(w, hb[kb], c);
        }
        ga[mt] = to_bf(tanh4(a));
        gb[mt] = to_bf(tanh4(c));
    }
    ya = B[4 * 8 + q];
    yb = ya;
    for (int kb = 0; kb < 4; kb++) {
        const s16x4 w = W[(20 + kb) * 64 + lane];
        ya = mfma(w, ga[kb], ya);
        yb = mfma(w, gb[kb], yb);
    }
}

template <class P>
struct ZMlp {
    static constexpr int gcd64(int t) { return t % 64 == 0 ? 64 : t % 32 == 0 ? 32 : t % 16 == 0 ? 16 : t % 8 == 0 ? 8 : t % 4 == 0 ? 4 : t % 2 == 0 ? 2 : 1; }
    static constexpr int LPB = 64 / gcd64(P::TPL);  // P::TPL * LPB is a multiple of 64
    static constexpr int CS = LPB + 1;
    static constexpr int S = LPB * P::N;            // samples of a block (a multiple of 32 for N % 4 == 0)
    static constexpr int GS = S + 16;               // row stride of the sample array
    static constexpr int THREADS = 8 * P::TPL * LPB;  // 8 groups
    static constexpr long LDS = 4L * (2 * 8 * P::N * CS + 9 * GS) + 8L * 24 * 64 + 16L * 36;
    // two blocks per CU (160 KiB of LDS)
    static constexpr bool OK = P::N % 4 == 0 && LDS <= 80 * 1024 && THREADS <= 1024;
};

template <class P>
__global__ __launch_bounds__(ZMlp<P>::THREADS) void k_les_z_mlp(const float2* __restrict__ Z, float2* __restrict__ T,
                                                                const float2* __restrict__ tw, int n,
                                                                const float* __restrict__ Pm,
                                                                const float* __restrict__ sc, int clip,
                                                                float smag) {
    using Z_ = ZMlp<P>;
    constexpr int N = P::N, LPB = Z_::LPB, CS = Z_::CS, S = Z_::S, GS = Z_::GS;
    constexpr int GT = P::TPL * LPB;  // threads per group (64)
    __shared__ float sre[8 * N * CS];
    __shared__ float sim[8 * N * CS];
    __shared__ float G[9 * GS];
    __shared__ s16x4 Wop[24 * 64];
    __shared__ f32x4 Bop[9 * 4];
    const int g = threadIdx.x / GT;
    const int c = (threadIdx.x % GT) % LPB, tid = (threadIdx.x % GT) / LPB;
    const int tiles = (n + LPB - 1) / LPB;
    const int j = blockIdx.x / tiles;
    const int i = (blockIdx.x % tiles) * LPB + c;
    const bool valid = i < n && j < n;
    const long N3 = (long)n * n * n, es = (long)n * n;
    const long cb = valid ? (long)i + (long)j * n : 0;
    float* gre = sre + g * N * CS;
    float* gim = sim + g * N * CS;
    auto idx = [&](int pos) { return pos * CS + c; };
    // operand o of lane l at Wop[64 o + l], as load_fwd forms them: W1^T tile
    // mt at o = mt, W2^T tile (mt, kb) at 4 + 4 mt + kb, W3^T tile kb at 20 + kb
    for (int e = threadIdx.x; e < 24 * 64; e += blockDim.x) {
        const int o = e / 64, l = e % 64, r = l % 16, q = l / 16;
        float v[4];
        for (int jj = 0; jj < 4; jj++) {
            if (o < 4)
                v[jj] = Pm[OFF_W1 + (4 * q + jj) * 64 + 16 * o + r];
            else if (o < 20)
                v[jj] = Pm[OFF_W2 + (16 * ((o - 4) % 4) + 4 * q + jj) * 64 + 16 * ((o - 4) / 4) + r];
            else
                v[jj] = Pm[OFF_W3 + (16 * (o - 20) + 4 * q + jj) * 16 + r];
        }
        Wop[e] = pack4(v[0], v[1], v[2], v[3]);
    }
    // bias tile k (b1 tiles 0-3, b2 tiles 4-7, b3 8) of lane group q at Bop[4 k + q]
    if (threadIdx.x < 36) {
        const int k = threadIdx.x / 4, q = threadIdx.x % 4;
        const int base = k < 4 ? OFF_B1 + 16 * k : k < 8 ? OFF_B2 + 16 * (k - 4) : OFF_B3;
        f32x4 b;
        for (int jj = 0; jj < 4; jj++) b[jj] = Pm[base + 4 * q + jj];
        Bop[threadIdx.x] = b;
    }
    {  // inverse z transforms of the gradients -> samples
        float2 x[P::E];
        for_in<P>(tid, [&](int s, int pos) { x[s] = g < 5 ? Z[g * N3 + cb + pos * es] : make_float2(0.f, 0.f); });
        line_fft<P, 1, float>(x, tid, gre, gim, idx, tw);
        const float scale = (float)(1.0 / (double)N3);
        if (g < 5)
            for_out<P>(tid, [&](int s, int pos) {
                const float2 v = rmul<float>(scale, x[s]);
                G[(2 * g) * GS + pos * LPB + c] = v.x;
                if (g < 4) G[(2 * g + 1) * GS + pos * LPB + c] = v.y;
            });
    }
    __syncthreads();
    {  // the network, two 16-sample tiles per trip and wave; stresses over rows 0-5
        const int lane = threadIdx.x % 64, r = lane % 16, q = lane / 16;
        const int wave = threadIdx.x / 64, nw = blockDim.x / 64;
        const f32x4 sc_in = load_scale_in(sc, lane);
        f32x4 sc_out;
        for (int jj = 0; jj < 4; jj++) sc_out[jj] = 4 * q + jj < kOut ? sc[kIn + 4 * q + jj] : 0.f;
        auto raw = [&](int s0) {
            f32x4 xf;
            for (int jj = 0; jj < 4; jj++) xf[jj] = 4 * q + jj < kIn ? G[(4 * q + jj) * GS + s0 + r] : 0.f;
            return xf;
        };
        for (int s00 = wave * 32; s00 < S; s00 += nw * 32) {
            const f32x4 ra = raw(s00), rb = raw(s00 + 16);
            f32x4 ya, yb;
            // the operands are read from LDS at each use (not held in ~84
            // VGPRs): the laundered lane index keeps the reads in the loop
            int ln = lane;
            asm volatile("" : "+v"(ln));
            forward2_lds(Wop, Bop, ln, q, to_bf(ra * sc_in), to_bf(rb * sc_in), ya, yb);
            for (int half = 0; half < 2; half++) {
                f32x4 y = (half ? yb : ya) * sc_out;
                if (clip) y = clip_backscatter(y, half ? rb : ra, r);
                if (smag > 0.f) y = smagorinsky_add(y, half ? rb : ra, r, smag);
                if (q < 2)
                    for (int jj = 0; jj < 4; jj++)
                        if (4 * q + jj < kOut) G[(4 * q + jj) * GS + s00 + 16 * half + r] = y[jj];
            }
        }
    }
    __syncthreads();
    {  // forward z transforms of tau_{2g} + i tau_{2g+1}, g < 3
        float2 x[P::E];
        for_in<P>(tid, [&](int s, int pos) {
            const int sm = pos * LPB + c;
            x[s] = g < 3 ? make_float2(G[(2 * g) * GS + sm], G[(2 * g + 1) * GS + sm]) : make_float2(0.f, 0.f);
        });
        line_fft<P, -1, float>(x, tid, gre, gim, idx, tw);
        if (valid && g < 3) for_out<P>(tid, [&](int s, int pos) { T[g * N3 + cb + pos * es] = x[s]; });
    }
}

__global__ void k_copy_c(const float2* __restrict__ a, float2* __restrict__ b, long n) {
    for (long p = blockIdx.x * (long)blockDim.x + threadIdx.x; p < n; p += (long)gridDim.x * blockDim.x) b[p] = a[p];
}

// ---------------------------------------------------------------- force-matching loss
// Residual of the closure force against a target, r = M P (f - ft) in place
// of f (P: divergence-free projection, M: zero on the Nyquist planes and, with
// dealias, where some |k_i| > n/3), with per-block sums of |r|^2 (FP64).
constexpr int kResidBlocks = 512;

// With kw2_inv = 1/k_w^2 > 0 each mode is weighted by w(k) = 1/(1 + |k|^2/k_w^2):
// the sums hold w|r|^2 and f is left holding w r, the seed of the adjoint.
__global__ __launch_bounds__(256) void k_force_resid(float2* __restrict__ f, const float2* __restrict__ ft, long fs,
                                                     int n, int dealias, float kw2_inv, double* __restrict__ part) {
    const long N = (long)n * n * n;
    double acc = 0;
    for (long p = blockIdx.x * (long)blockDim.x + threadIdx.x; p < N; p += (long)gridDim.x * blockDim.x) {
        const int i = (int)(p % n), j = (int)((p / n) % n), k = (int)(p / ((long)n * n));
        const float kk[3] = {kd(i, n), kd(j, n), kd(k, n)};
        bool zero = 2 * i == n || 2 * j == n || 2 * k == n || (i == 0 && j == 0 && k == 0);
        if (dealias == 1) zero = zero || 3 * fabsf(kk[0]) > n || 3 * fabsf(kk[1]) > n || 3 * fabsf(kk[2]) > n;
        // mode 2: the residual is the solver's filtered force, so its square carries sigma^2
        const float sig = dealias == 2 ? expfilter_f(kk[0], n) * expfilter_f(kk[1], n) * expfilter_f(kk[2], n) : 1.f;
        float2 d[3];
        for (int a = 0; a < 3; a++) {
            const float2 x = f[a * fs + p], y = ft[a * fs + p];
            d[a] = make_float2(x.x - y.x, x.y - y.y);
        }
        float wk = 1.f;
        if (zero) {
            for (int a = 0; a < 3; a++) d[a] = make_float2(0.f, 0.f);
        } else {
            const float k2 = kk[0] * kk[0] + kk[1] * kk[1] + kk[2] * kk[2];
            const float dr = (kk[0] * d[0].x + kk[1] * d[1].x + kk[2] * d[2].x) / k2;
            const float di = (kk[0] * d[0].y + kk[1] * d[1].y + kk[2] * d[2].y) / k2;
            for (int a = 0; a < 3; a++) d[a] = make_float2(d[a].x - kk[a] * dr, d[a].y - kk[a] * di);
            wk = sig * sig / (1.f + k2 * kw2_inv);
        }
        for (int a = 0; a < 3; a++) {
            f[a * fs + p] = make_float2(wk * d[a].x, wk * d[a].y);
            acc += (double)wk * ((double)d[a].x * d[a].x + (double)d[a].y * d[a].y);
        }
    }
    __shared__ double red[256];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

__global__ void k_sum_partials(const double* __restrict__ part, int np, double scale, double* __restrict__ out) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        double s = 0;
        for (int i = 0; i < np; i++) s += part[i];
        *out = s * scale;
    }
}

// dL/dy (6 planar fields) from the packed physical gradients of the residual,
// g_{3a+b} = d r_a / d x_b: the adjoint of f = -div tau applied to 2 r / N,
// times the output scale
__global__ void k_force_dy(const float2* __restrict__ Z, long zs, float* __restrict__ D, long ld, long n, float c,
                           const float* __restrict__ sc_out) {
    for (long p = blockIdx.x * (long)blockDim.x + threadIdx.x; p < n; p += (long)gridDim.x * blockDim.x) {
        float g[10];
        for (int m = 0; m < 5; m++) {
            const float2 z = Z[m * zs + p];
            g[2 * m] = z.x;
            g[2 * m + 1] = z.y;
        }
        const float d[6] = {g[0], g[4], g[8], g[1] + g[3], g[2] + g[6], g[5] + g[7]};
        for (int o = 0; o < 6; o++) D[o * ld + p] = c * sc_out[o] * d[o];
    }
}

__global__ void k_real_to_c(const float* __restrict__ u, float2* __restrict__ o, long n) {
    for (long p = blockIdx.x * (long)blockDim.x + threadIdx.x; p < n; p += (long)gridDim.x * blockDim.x)
        o[p] = make_float2(u[p], 0.f);
}

__global__ void k_real_to_z(const float* __restrict__ u, double2* __restrict__ o, long n) {
    for (long p = blockIdx.x * (long)blockDim.x + threadIdx.x; p < n; p += (long)gridDim.x * blockDim.x)
        o[p] = make_double2((double)u[p], 0.0);
}

__global__ void k_z_to_c2(const double2* __restrict__ a, float2* __restrict__ b, long n) {
    for (long p = blockIdx.x * (long)blockDim.x + threadIdx.x; p < n; p += (long)gridDim.x * blockDim.x)
        b[p] = make_float2((float)a[p].x, (float)a[p].y);
}

// zero the Nyquist modes, then V <- V - k (k.V) / |k|^2 (FP64)
__global__ void k_les_project(double2* __restrict__ V, long fs, int n) {
    const long N = (long)n * n * n;
    for (long p = blockIdx.x * (long)blockDim.x + threadIdx.x; p < N; p += (long)gridDim.x * blockDim.x) {
        const int i = (int)(p % n), j = (int)((p / n) % n), k = (int)(p / ((long)n * n));
        if (2 * i == n || 2 * j == n || 2 * k == n) {
            for (int a = 0; a < 3; a++) V[a * fs + p] = make_double2(0.0, 0.0);
            continue;
        }
        const double kk[3] = {(double)(i <= n / 2 ? i : i - n), (double)(j <= n / 2 ? j : j - n),
                              (double)(k <= n / 2 ? k : k - n)};
        const double k2 = kk[0] * kk[0] + kk[1] * kk[1] + kk[2] * kk[2];
        if (k2 == 0) continue;
        double2 v[3];
        double dr = 0, di = 0;
        for (int a = 0; a < 3; a++) {
            v[a] = V[a * fs + p];
            dr += kk[a] * v[a].x;
            di += kk[a] * v[a].y;
        }
        for (int a = 0; a < 3; a++)
            V[a * fs + p] = make_double2(v[a].x - kk[a] * dr / k2, v[a].y - kk[a] * di / k2);
    }
}

// X[f*ldx + s] = component f % 2 of Z[f / 2]
__global__ void k_unpack_planar(const float2* __restrict__ Z, long zs, float* __restrict__ X, long ldx, long n) {
    for (long p = blockIdx.x * (long)blockDim.x + threadIdx.x; p < n; p += (long)gridDim.x * blockDim.x)
        for (int f = 0; f < 9; f++) {
            const float2 z = Z[(f >> 1) * zs + p];
            X[f * ldx + p] = (f & 1) ? z.y : z.x;
        }
}

static unsigned lgrid(long n) {
    long g = (n + 255) / 256;
    if (g > 4096) g = 4096;
    return (unsigned)(g < 1 ? 1 : g);
}

#define RC(x)                \
    do {                     \
        int _rc = (x);       \
        if (_rc) return _rc; \
    } while (0)

template <class P, int LPB>
int launch_grad_x(const float2* V, float2* Z, int n, int half, hipStream_t st) {
    const float2* tw = (const float2*)twiddles(P::N, SH_F32);
    if (!tw) return SH_ENOMEM;
    const long nlines = (long)n * n;
    hipLaunchKernelGGL((k_les_grad_x<P, LPB>), dim3((unsigned)((nlines + LPB - 1) / LPB), 5), dim3(P::TPL * LPB), 0,
                       st, V, Z, tw, n, half);
    SH_LAUNCH_CHECK("k_les_grad_x");
    return SH_OK;
}

// the fused derivative + x pass of the gradient fields (reads V, writes Z)
int les_grad_x(const float2* V, float2* Z, int n, int half, hipStream_t st) {
    int rc = SH_EUNSUPPORTED;
    switch (n) {
#define SH_CASE(n_, P_, L_, S_)                   \
    case n_:                                      \
        rc = launch_grad_x<P_, L_>(V, Z, n, half, st); \
        break;
        SH_FOR_EACH_FFT_SIZE(SH_CASE)
#undef SH_CASE
    }
    if (rc == SH_EUNSUPPORTED) set_error("les: unsupported grid size");
    return rc;
}

template <class P>
int launch_grad_xy(const float2* V, float2* Z, int n, int half, hipStream_t st) {
    constexpr bool OK = 2L * 4 * P::N * (P::N + P::N / 16 + 1) <= 65536 && P::TPL * P::N <= 1024;
    if constexpr (!OK) {
        return SH_EUNSUPPORTED;
    } else {
        const float2* tw = (const float2*)twiddles(P::N, SH_F32);
        if (!tw) return SH_ENOMEM;
        hipLaunchKernelGGL((k_les_grad_xy<P>), dim3((unsigned)n, 5), dim3(P::TPL * P::N), 0, st, V, Z, tw, n, half);
        SH_LAUNCH_CHECK("k_les_grad_xy");
        return SH_OK;
    }
}

// the gradient fields after their x and y passes: one plane kernel where the
// plane fits in LDS (n <= 64), else the x pass and the y pass
int les_grad_xy(const float2* V, float2* Z, int n, int half, hipStream_t st) {
    int rc = SH_EUNSUPPORTED;
    switch (n) {
#define SH_CASE(n_, P_, L_, S_)                               \
    case n_:                                                  \
        rc = launch_grad_xy<P_>(V, Z, n, half, st);           \
        break;
        SH_FOR_EACH_FFT_SIZE(SH_CASE)
#undef SH_CASE
    }
    if (rc != SH_EUNSUPPORTED) return rc;
    RC(les_grad_x(V, Z, n, half, st));
    const long sxy = (long)n * n;
    Epi<float> plain{1.f, 0, 1, -1, -1};
    return pass_strided<float>(Z, n, n, 1, n, sxy, n, 5, sxy * n, +1, plain, st, ROLE_Y);
}

// physical packed gradient fields Z (5 complex64 fields) of V (3 Fourier
// fields, n^3 each; half: V in the half-spectrum layout)
int les_gradients(const float2* V, float2* Z, int n, int half, hipStream_t st) {
    ProfScope ps("les_fft_inv", st);
    RC(les_grad_xy(V, Z, n, half, st));
    // the z pass, scaled by 1/n^3
    const long sxy = (long)n * n, N = sxy * n;
    Epi<float> last{(float)(1.0 / (double)N), 1, 1, -1, -1};
    return pass_strided<float>(Z, n, n, 1, n, n, sxy, 5, N, +1, last, st, ROLE_Z);
}

template <class P>
int launch_x_div(const float2* T, float2* F, float2* H, int n, int mode, hipStream_t st) {
    constexpr int PPB = 256 / (6 * P::TPL) > 0 ? 256 / (6 * P::TPL) : 1;
    const float2* tw = (const float2*)twiddles(P::N, SH_F32);
    if (!tw) return SH_ENOMEM;
    const long npairs = (long)n * n / 2 + 2;
    hipLaunchKernelGGL((k_les_x_div<P, PPB>), dim3((unsigned)((npairs + PPB - 1) / PPB)), dim3(P::TPL * 6 * PPB), 0,
                       st, T, F, H, tw, n, mode);
    SH_LAUNCH_CHECK("k_les_x_div");
    return SH_OK;
}

template <class P>
int launch_z_mlp(const float2* Z, float2* T, int n, const float* Pm, const float* sc, int clip, float smag,
                 hipStream_t st) {
    if constexpr (!ZMlp<P>::OK) {
        return SH_EUNSUPPORTED;
    } else {
        const float2* tw = (const float2*)twiddles(P::N, SH_F32);
        if (!tw) return SH_ENOMEM;
        const int tiles = (n + ZMlp<P>::LPB - 1) / ZMlp<P>::LPB;
        hipLaunchKernelGGL((k_les_z_mlp<P>), dim3((unsigned)(tiles * n)), dim3(ZMlp<P>::THREADS), 0, st, Z, T, tw,
                           n, Pm, sc, clip, smag);
        SH_LAUNCH_CHECK("k_les_z_mlp");
        return SH_OK;
    }
}

// gradients after their x and y passes -> packed stresses after their z
// pass: one fused kernel where its LDS fits (n = 8, 16, 32, 64; Z is only
// read), else the z pass (in place in Z), the network and the z pass as
// separate launches
int les_z_mlp(sh_mlp_t m, float2* Z, float2* T, int n, int clip, float smag, hipStream_t st) {
    int rc = SH_EUNSUPPORTED;
    switch (n) {
#define SH_CASE(n_, P_, L_, S_)                                                       \
    case n_:                                                                          \
        rc = launch_z_mlp<P_>(Z, T, n, mlp_params(m), mlp_scales(m), clip, smag, st); \
        break;
        SH_FOR_EACH_FFT_SIZE(SH_CASE)
#undef SH_CASE
    }
    if (rc != SH_EUNSUPPORTED) return rc;
    const long sxy = (long)n * n, N = sxy * n;
    Epi<float> last{(float)(1.0 / (double)N), 1, 1, -1, -1};
    RC(pass_strided<float>(Z, n, n, 1, n, n, sxy, 5, N, +1, last, st, ROLE_Z));
    RC(mlp_infer_packed(m, Z, N, N, T, N, clip, smag, st));
    Epi<float> plain{1.f, 0, 1, -1, -1};
    return pass_strided<float>(T, n, n, 1, n, n, sxy, 3, N, -1, plain, st, ROLE_Z);
}

// f = -div tau from the packed stresses T after their z pass (3 fields,
// overwritten): the forward y pass, then the x pass fused with the Hermitian
// unpacking and the divergence (k_les_x_div, `mode` as there)
int les_force_yx(float2* T, float2* F, float2* H, int n, int mode, hipStream_t st) {
    const long sxy = (long)n * n, N = sxy * n;
    Epi<float> plain{1.f, 0, 1, -1, -1};
    RC(pass_strided<float>(T, n, n, 1, n, sxy, n, 3, N, -1, plain, st, ROLE_Y));
    int rc = SH_EUNSUPPORTED;
    switch (n) {
#define SH_CASE(n_, P_, L_, S_)                                 \
    case n_:                                                    \
        rc = launch_x_div<P_>(T, F, H, n, mode, st);           \
        break;
        SH_FOR_EACH_FFT_SIZE(SH_CASE)
#undef SH_CASE
    }
    if (rc == SH_EUNSUPPORTED) set_error("les: unsupported grid size");
    return rc;
}

}  // namespace sh

using namespace sh;

struct sh_les_s {
    int n;
    long N;
    sh_mlp_t model;
    int clip;
    float smag;  // (Cs Delta)^2 of the Smagorinsky stress added to the network's, 0: none
    hipStream_t st;
    sh_ns_t ns;
    // force slots (3 fields each; slot 0 is the one the solver reads between
    // calls), packed gradients (5), packed stresses (3), FP32 velocity (3),
    // the previous step's force for the lagged closure (3)
    float2 *F, *F1, *Z, *T, *Vc, *H;
    double2* V64;  // 3 fields for the initial condition
    int lag;       // 1: the closure of step t forces step t+1 (extrapolated)
    bool started;  // an initial condition has been set
    // the closure runs on `side` while the solver's kernels run on st; with
    // the synchronous closure kernel E waits for ev_force, with the lagged
    // one only for ev_read (the closure has read V); the next closure waits
    // for ev_state
    hipStream_t side;
    hipEvent_t ev_state, ev_force, ev_read;
};

namespace {

// 1 when the solver steps the half-spectrum layout (its velocity holds the
// planes kz < n/2 only)
int les_half(sh_les_s* h) {
    int half = 0;
    return sh_ns_get_layout(h->ns, &half) == SH_OK && half ? 1 : 0;
}

// f = -div tau(V) into F, every plane (mode and H as k_les_x_div); half: V is
// in the half-spectrum layout; ev_read, if given, is recorded once V has
// been read
int sgs_force(sh_les_s* h, const float2* V, int half, float2* F, float2* H, int mode, hipStream_t st,
              hipEvent_t ev_read) {
    const int n = h->n;
    {
        ProfScope ps("les_fft_inv", st);
        RC(les_grad_xy(V, h->Z, n, half, st));
        if (ev_read) SH_HIP(hipEventRecord(ev_read, st));
    }
    {
        ProfScope ps("les_mlp", st);
        RC(les_z_mlp(h->model, h->Z, h->T, n, h->clip, h->smag, st));
    }
    ProfScope ps("les_fft_fwd", st);
    return les_force_yx(h->T, F, H, n, mode, st);
}

// one LES step with the synchronous closure: the closure of V^t on the side
// stream, concurrent with the solver's kernels A-D (which read V and the work
// fields and write only the latter); kernel E reads the force and then
// overwrites V, after which the next closure may start
int les_step(sh_les_s* h) {
    if (!h->model) return sh_ns_steps(h->ns, 1);
    const float2* V = (const float2*)ns_fused_velocity(h->ns);
    SH_HIP(hipEventRecord(h->ev_state, h->st));
    SH_HIP(hipStreamWaitEvent(h->side, h->ev_state, 0));
    RC(sgs_force(h, V, les_half(h), h->F, nullptr, 0, h->side, nullptr));
    SH_HIP(hipEventRecord(h->ev_force, h->side));
    RC(ns_step_phase(h->ns, 0));
    SH_HIP(hipStreamWaitEvent(h->st, h->ev_force, 0));
    return ns_step_phase(h->ns, 1);
}

// nsteps steps with the lagged closure.  Step t reads the force slot t % 2,
// which holds 2 f(V^{t-1}) - f(V^{t-2}) (second order in time, like the AB2
// term it is added to); the closure of V^t runs on the side stream
// concurrently with the solver's whole step t and writes slot (t+1) % 2.
// Kernel E of step t waits only for the closure's first kernel, the reader
// of V.  The call ends with the side stream joined and the pending force in
// slot 0, so any number of steps per call (and graph replays of a call)
// continue the same sequence.
int les_steps_lagged(sh_les_s* h, int nsteps) {
    float2* slot[2] = {h->F, h->F1};
    const int half = les_half(h);
    for (int t = 0; t < nsteps; t++) {
        const float2* V = (const float2*)ns_fused_velocity(h->ns);
        SH_HIP(hipEventRecord(h->ev_state, h->st));
        SH_HIP(hipStreamWaitEvent(h->side, h->ev_state, 0));
        RC(sgs_force(h, V, half, slot[(t + 1) % 2], h->H, 2, h->side, h->ev_read));
        RC(sh_ns_set_forcing(h->ns, slot[t % 2]));
        RC(ns_step_phase(h->ns, 0));
        SH_HIP(hipStreamWaitEvent(h->st, h->ev_read, 0));
        RC(ns_step_phase(h->ns, 1));
    }
    SH_HIP(hipEventRecord(h->ev_force, h->side));
    SH_HIP(hipStreamWaitEvent(h->st, h->ev_force, 0));
    if (nsteps % 2) {
        hipLaunchKernelGGL(k_copy_c, dim3(lgrid(3 * h->N)), dim3(256), 0, h->st, h->F1, h->F, 3 * h->N);
        SH_LAUNCH_CHECK("k_copy_c");
    }
    return sh_ns_set_forcing(h->ns, h->F);
}

int les_run(sh_les_s* h, int nsteps) {
    if (h->model && h->lag) return les_steps_lagged(h, nsteps);
    for (int t = 0; t < nsteps; t++) RC(les_step(h));
    return SH_OK;
}

// a timing unit for the solver's schedule tuner: one synchronous step, or two
// lagged ones (the lagged call ends with a join)
int les_step_cb(void* p) {
    sh_les_s* h = (sh_les_s*)p;
    return les_run(h, h->model && h->lag ? 2 : 1);
}

void free_les(sh_les_s* h) {
    if (h->ev_state) (void)hipEventDestroy(h->ev_state);
    if (h->ev_force) (void)hipEventDestroy(h->ev_force);
    if (h->ev_read) (void)hipEventDestroy(h->ev_read);
    if (h->side) (void)hipStreamDestroy(h->side);
    if (h->ns) sh_ns_destroy(h->ns);
    float2* bufs[6] = {h->F, h->F1, h->Z, h->T, h->Vc, h->H};
    for (float2* b : bufs)
        if (b) pool_free(b);
    if (h->V64) pool_free(h->V64);
    delete h;
}

}  // namespace

extern "C" {

int sh_les_create(sh_les_t* handle, int n, double nu, double K, sh_mlp_t model, int clip, void* stream) {
    if (!handle || n <= 0) {
        set_error("sh_les_create: invalid argument");
        return SH_EINVAL;
    }
    RC(ensure_init());
    sh_les_s* h = new sh_les_s();
    h->n = n;
    h->N = (long)n * n * n;
    h->model = model;
    h->clip = clip != 0;
    h->st = as_stream(stream);
    // the solver steps the half-spectrum layout once its start is real; the
    // closure reads its velocity in place in either layout
    int rc = sh_ns_create(&h->ns, n, n, n, nu, K, SH_F32, 1, stream);
    // the closure reads V1 in place every step: no merged AB2 history
    if (!rc) ns_keep_velocity(h->ns);
    void* p;
    const long sizes[6] = {3, 3, 5, 3, 3, 3};
    float2** dst[6] = {&h->F, &h->F1, &h->Z, &h->T, &h->Vc, &h->H};
    for (int i = 0; i < 6; i++) {
        p = nullptr;
        if (!rc) rc = pool_alloc(&p, sizeof(float2) * h->N * sizes[i]);
        *dst[i] = (float2*)p;
    }
    p = nullptr;
    if (!rc) rc = pool_alloc(&p, sizeof(double2) * h->N * 3);
    h->V64 = (double2*)p;
    if (!rc && model) rc = sh_ns_set_forcing(h->ns, h->F);
    // default-priority side stream and ordering events (no timing)
    if (!rc && hipStreamCreateWithFlags(&h->side, hipStreamNonBlocking) != hipSuccess) rc = SH_EHIP;
    if (!rc && hipEventCreateWithFlags(&h->ev_state, hipEventDisableTiming) != hipSuccess) rc = SH_EHIP;
    if (!rc && hipEventCreateWithFlags(&h->ev_force, hipEventDisableTiming) != hipSuccess) rc = SH_EHIP;
    if (!rc && hipEventCreateWithFlags(&h->ev_read, hipEventDisableTiming) != hipSuccess) rc = SH_EHIP;
    if (rc) {
        if (rc == SH_EHIP) set_error("sh_les_create: stream or event creation failed");
        free_les(h);
        return rc;
    }
    *handle = h;
    return SH_OK;
}

int sh_les_destroy(sh_les_t h) {
    if (!h) return SH_OK;
    (void)hipStreamSynchronize(h->st);
    (void)hipStreamSynchronize(h->side);
    free_les(h);
    return SH_OK;
}

int sh_les_set_initial_physical(sh_les_t h, const void* ubar3) {
    if (!h || !ubar3) return SH_EINVAL;
    const long N = h->N;
    hipLaunchKernelGGL(k_real_to_z, dim3(lgrid(3 * N)), dim3(256), 0, h->st, (const float*)ubar3, h->V64, 3 * N);
    SH_LAUNCH_CHECK("k_real_to_z");
    RC(fft3d<double>(h->V64, 3, N, h->n, h->n, h->n, -1, 0, h->st));
    hipLaunchKernelGGL(k_les_project, dim3(lgrid(N)), dim3(256), 0, h->st, h->V64, N, h->n);
    SH_LAUNCH_CHECK("k_les_project");
    RC(sh_ns_set_initial(h->ns, h->V64));
    if (h->model) {
        hipLaunchKernelGGL(k_z_to_c2, dim3(lgrid(3 * N)), dim3(256), 0, h->st, h->V64, h->Vc, 3 * N);
        SH_LAUNCH_CHECK("k_z_to_c2");
        // the force of the start (the lagged closure's history starts there too)
        RC(sgs_force(h, h->Vc, 0, h->F, h->H, 1, h->st, nullptr));
    }
    h->started = true;
    return sh_ns_startup(h->ns);
}

int sh_les_steps(sh_les_t h, int nsteps) {
    if (!h || nsteps < 0) return SH_EINVAL;
    return les_run(h, nsteps);
}

int sh_les_set_smagorinsky(sh_les_t h, double cs) {
    if (!h || !(cs >= 0) || !h->model) {
        set_error("sh_les_set_smagorinsky: invalid argument (an LES with a model, Cs >= 0)");
        return SH_EINVAL;
    }
    const double delta = 2.0 * M_PI / h->n;  // the coarse grid spacing, the filter width
    h->smag = (float)((cs * delta) * (cs * delta));
    return SH_OK;
}

int sh_les_set_lag(sh_les_t h, int lag) {
    if (!h || lag < 0 || lag > 1) {
        set_error("sh_les_set_lag: lag must be 0 (synchronous closure) or 1 (lagged, extrapolated)");
        return SH_EINVAL;
    }
    if (lag && !h->lag && h->model && h->started) {
        // restart the history at the current state: the next step is forced
        // by f(V), as after a start
        RC(sgs_force(h, (const float2*)ns_fused_velocity(h->ns), les_half(h), h->F, h->H, 1, h->st, nullptr));
    }
    h->lag = lag;
    return sh_ns_set_forcing(h->ns, h->model ? h->F : nullptr);
}

int sh_les_set_dealias(sh_les_t h, int mode) {
    if (!h) return SH_EINVAL;
    return sh_ns_set_dealias(h->ns, mode);
}

int sh_les_autotune(sh_les_t h, int reps) {
    if (!h || reps <= 0) return SH_EINVAL;
    RC(sh_autotune_fft3d(h->n, h->n, h->n, SH_F32, reps, h->st));
    // the solver's schedules timed on the whole LES step (closure overlapped
    // on the side stream), which is what they run in.  The solver restores
    // its own state; the lagged closure's pending force and history are
    // saved and restored here.
    const bool keep = h->model && h->lag;
    void* save = nullptr;
    const size_t fb = sizeof(float2) * 3 * h->N;
    if (keep) {
        RC(pool_alloc(&save, 2 * fb));
        SH_HIP(hipMemcpyAsync(save, h->F, fb, hipMemcpyDeviceToDevice, h->st));
        SH_HIP(hipMemcpyAsync((char*)save + fb, h->H, fb, hipMemcpyDeviceToDevice, h->st));
    }
    int rc = ns_autotune_with(h->ns, reps, &les_step_cb, h);
    if (keep) {
        if (hipMemcpyAsync(h->F, save, fb, hipMemcpyDeviceToDevice, h->st) != hipSuccess ||
            hipMemcpyAsync(h->H, (char*)save + fb, fb, hipMemcpyDeviceToDevice, h->st) != hipSuccess ||
            hipStreamSynchronize(h->st) != hipSuccess) {
            set_error("sh_les_autotune: restoring the closure state failed");
            if (!rc) rc = SH_EHIP;
        }
        pool_free(save);
    }
    return rc;
}

int sh_les_get_velocity(sh_les_t h, void* out_v3) {
    if (!h || !out_v3) return SH_EINVAL;
    return sh_ns_get_velocity(h->ns, out_v3);
}

int sh_les_stress(sh_les_t h, void* out_tau6) {
    if (!h || !out_tau6 || !h->model) {
        set_error("sh_les_stress: invalid argument (an LES with a model)");
        return SH_EINVAL;
    }
    RC(les_gradients((const float2*)ns_fused_velocity(h->ns), h->Z, h->n, les_half(h), h->st));
    return mlp_infer_packed_planar(h->model, h->Z, h->N, h->N, (float*)out_tau6, h->N, h->clip, h->smag, h->st);
}

int sh_les_force(sh_les_t h, void* out_f3) {
    if (!h || !out_f3 || !h->model) {
        set_error("sh_les_force: invalid argument (an LES with a model)");
        return SH_EINVAL;
    }
    // straight into the caller's buffer: the force slots hold the lagged
    // closure's pending force
    return sgs_force(h, (const float2*)ns_fused_velocity(h->ns), les_half(h), (float2*)out_f3, nullptr, 0, h->st,
                     nullptr);
}

int sh_les_pending_force(sh_les_t h, void* out_f3) {
    if (!h || !out_f3 || !h->model) {
        set_error("sh_les_pending_force: invalid argument (an LES with a model)");
        return SH_EINVAL;
    }
    SH_HIP(hipMemcpyAsync(out_f3, h->F, sizeof(float2) * 3 * h->N, hipMemcpyDeviceToDevice, h->st));
    return SH_OK;
}

int sh_les_features(int n, const void* ubar3, void* X, long ldx, void* stream) {
    if (n <= 0 || !ubar3 || !X || ldx < (long)n * n * n) {
        set_error("sh_les_features: invalid argument");
        return SH_EINVAL;
    }
    RC(ensure_init());
    hipStream_t st = as_stream(stream);
    const long N = (long)n * n * n;
    void *v = nullptr, *z = nullptr;
    int rc = pool_alloc(&v, sizeof(float2) * 3 * N);
    if (!rc) rc = pool_alloc(&z, sizeof(float2) * 5 * N);
    if (!rc) {
        hipLaunchKernelGGL(k_real_to_c, dim3(lgrid(3 * N)), dim3(256), 0, st, (const float*)ubar3, (float2*)v, 3 * N);
        rc = hipGetLastError() == hipSuccess ? SH_OK : SH_EHIP;
    }
    if (!rc) rc = fft3d<float>(v, 3, N, n, n, n, -1, 0, st);
    if (!rc) rc = les_gradients((const float2*)v, (float2*)z, n, 0, st);
    if (!rc) {
        hipLaunchKernelGGL(k_unpack_planar, dim3(lgrid(N)), dim3(256), 0, st, (const float2*)z, N, (float*)X, ldx, N);
        rc = hipGetLastError() == hipSuccess ? SH_OK : SH_EHIP;
    }
    if (hipStreamSynchronize(st) != hipSuccess && !rc) rc = SH_EHIP;
    if (v) pool_free(v);
    if (z) pool_free(z);
    if (rc) set_error("sh_les_features: failed");
    return rc;
}

}  // extern "C"

namespace {

// the force-matching loss and its parameter gradient (into the handle's
// gradient buffer); the loss itself is left in part[kResidBlocks]
int force_loss_grad(sh_mlp_t m, int n, const float* X, const float2* ft, int dealias, double weight, double kw,
                    double** loss_dev) {
    const long N = (long)n * n * n;
    const size_t bytes = sizeof(float2) * N * (3 + 3 + 5) + sizeof(float) * N * 6 + sizeof(double) * (kResidBlocks + 1);
    char* base = (char*)mlp_aux(m, bytes);
    if (!base) return SH_ENOMEM;
    float2* T = (float2*)base;
    float2* R = T + 3 * N;
    float2* Z = R + 3 * N;
    float* D = (float*)(Z + 5 * N);
    double* part = (double*)(base + sizeof(float2) * N * 11 + sizeof(float) * N * 6);
    hipStream_t st = mlp_stream(m);
    // f = -div tau(X), as the LES forms it
    RC(mlp_infer_planar_packed(m, X, N, N, T, N, st));
    RC(fft3d<float>(T, 3, N, n, n, n, -1, 0, st));
    hipLaunchKernelGGL(k_les_div_unpack, dim3(lgrid(N)), dim3(256), 0, st, T, N, R, N, n);
    SH_LAUNCH_CHECK("k_les_div_unpack");
    // r = M P (f - ft) and |r|^2
    const float kw2_inv = kw > 0 ? (float)(1.0 / (kw * kw)) : 0.f;
    hipLaunchKernelGGL(k_force_resid, dim3(kResidBlocks), dim3(256), 0, st, R, ft, N, n, dealias, kw2_inv, part);
    SH_LAUNCH_CHECK("k_force_resid");
    const double Nd = (double)N;
    hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(64), 0, st, part, kResidBlocks, weight / (Nd * Nd),
                       part + kResidBlocks);
    SH_LAUNCH_CHECK("k_sum_partials");
    // dL/dtau by the adjoint of the divergence, then back through the network
    RC(les_gradients(R, Z, n, 0, st));
    hipLaunchKernelGGL(k_force_dy, dim3(lgrid(N)), dim3(256), 0, st, Z, N, D, N, N, (float)(2.0 * weight / Nd),
                       mlp_scales(m) + 9);
    SH_LAUNCH_CHECK("k_force_dy");
    RC(mlp_grad_dy(m, X, D, N, N, st));
    *loss_dev = part + kResidBlocks;
    return SH_OK;
}

int read_loss(sh_mlp_t m, const double* dev, double* loss) {
    if (!loss) return SH_OK;
    SH_HIP(hipMemcpyAsync(loss, dev, sizeof(double), hipMemcpyDeviceToHost, mlp_stream(m)));
    SH_HIP(hipStreamSynchronize(mlp_stream(m)));
    return SH_OK;
}

}  // namespace

extern "C" {

int sh_mlp_force_gradient(sh_mlp_t h, int n, const void* X, const void* ft, int dealias, double weight,
                          double k_weight, void* grad_out, double* loss) {
    if (!h || n <= 0 || !X || !ft || !(weight > 0) || k_weight < 0) {
        set_error("sh_mlp_force_gradient: invalid argument");
        return SH_EINVAL;
    }
    double* dev = nullptr;
    RC(force_loss_grad(h, n, (const float*)X, (const float2*)ft, dealias, weight, k_weight, &dev));
    if (grad_out) {
        int np = 0;
        RC(sh_mlp_param_count(&np));
        SH_HIP(hipMemcpyAsync(grad_out, mlp_grad_buffer(h), sizeof(float) * np, hipMemcpyDeviceToDevice,
                              mlp_stream(h)));
    }
    return read_loss(h, dev, loss);
}

int sh_mlp_force_train_step(sh_mlp_t h, int n, const void* X, const void* ft, int dealias, double weight,
                            double k_weight, float lr, double* loss) {
    if (!h || n <= 0 || !X || !ft || !(weight > 0) || k_weight < 0) {
        set_error("sh_mlp_force_train_step: invalid argument");
        return SH_EINVAL;
    }
    double* dev = nullptr;
    RC(force_loss_grad(h, n, (const float*)X, (const float2*)ft, dealias, weight, k_weight, &dev));
    RC(mlp_adam(h, lr, mlp_stream(h)));
    return read_loss(h, dev, loss);
}

}  // extern "C"
