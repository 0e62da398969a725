// mlp_core.hpp -- device pieces of the closure network (9 -> 64 -> 64 -> 6,
// tanh hidden layers) on BF16 MFMA with FP32 accumulation, shared by the
// network's own kernels (mlp.hip) and the fused LES z pass (les.hip).
//
// MFMA: v_mfma_f32_16x16x16_bf16 (__builtin_amdgcn_mfma_f32_16x16x16bf16_1k).
// Operand layout for lane l of the wavefront, r = l % 16, q = l / 16, j = 0..3:
//   A (16x16, M x K): A[r][4q + j]
//   B (16x16, K x N): B[4q + j][r]
//   D (16x16, M x N): D[4q + j][r]
// (checked by sh_mfma_probe).  The network runs "features x samples":
// Z^T = W^T X^T, so one layer's D tile (4 consecutive features of one sample
// per lane) is the next layer's B operand as it stands and inference needs no
// LDS.
//
// Packed parameters (floats): W1[16][64] (input rows 9..15 zero), b1[64],
// W2[64][64], b2[64], W3[64][16] (output columns 6..15 zero), b3[16], where
// W[i][o] maps input i to output o.
#pragma once

#include "sh_common.hpp"

namespace sh {

typedef short s16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef short s16x2 __attribute__((ext_vector_type(2)));

constexpr int kIn = 9, kOut = 6, kH = 64;
constexpr int OFF_W1 = 0, OFF_B1 = 16 * 64, OFF_W2 = OFF_B1 + 64, OFF_B2 = OFF_W2 + 64 * 64,
              OFF_W3 = OFF_B2 + 64, OFF_B3 = OFF_W3 + 64 * 16, NPARAM = OFF_B3 + 16;
constexpr int kGradWaves = 1024;  // partial-gradient slots

__device__ __forceinline__ f32x4 mfma(s16x4 a, s16x4 b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

typedef unsigned u32x2 __attribute__((ext_vector_type(2)));

// two FP32 -> BF16 conversions in one v_cvt_pk_bf16_f32 (round to nearest even),
// as the 32-bit register that holds them
__device__ __forceinline__ unsigned cvt2(float a, float b) {
    return __builtin_bit_cast(unsigned, __builtin_convertvector(f32x2{a, b}, bf16x2));
}

// the 4 x 16-bit MFMA operand is two such registers as they stand
__device__ __forceinline__ s16x4 pack4(float a, float b, float c, float d) {
    return __builtin_bit_cast(s16x4, u32x2{cvt2(a, b), cvt2(c, d)});
}

__device__ __forceinline__ s16x4 to_bf(f32x4 v) { return pack4(v[0], v[1], v[2], v[3]); }

// weights as MFMA A operands, rounded from the FP32 master parameters
struct FwdW {
    s16x4 w1t[4], w2t[4][4], w3t[4];  // W1^T tile mt; W2^T tile (mt, kb); W3^T tile kb
};
struct BwdW {
    s16x4 w3[4], w2[4][4];  // W3 tile mt (K = outputs); W2 tile (mt, kb)
};

__device__ void load_fwd(const float* __restrict__ P, FwdW& w, int lane) {
    const int r = lane % 16, q = lane / 16;
    for (int mt = 0; mt < 4; mt++) {
        float v[4];
        for (int j = 0; j < 4; j++) v[j] = P[OFF_W1 + (4 * q + j) * 64 + 16 * mt + r];
        w.w1t[mt] = pack4(v[0], v[1], v[2], v[3]);
        for (int kb = 0; kb < 4; kb++) {
            for (int j = 0; j < 4; j++) v[j] = P[OFF_W2 + (16 * kb + 4 * q + j) * 64 + 16 * mt + r];
            w.w2t[mt][kb] = pack4(v[0], v[1], v[2], v[3]);
        }
        for (int j = 0; j < 4; j++) v[j] = P[OFF_W3 + (16 * mt + 4 * q + j) * 16 + r];
        w.w3t[mt] = pack4(v[0], v[1], v[2], v[3]);
    }
}

__device__ void load_bwd(const float* __restrict__ P, BwdW& w, int lane) {
    const int r = lane % 16, q = lane / 16;
    for (int mt = 0; mt < 4; mt++) {
        float v[4];
        for (int j = 0; j < 4; j++) v[j] = P[OFF_W3 + (16 * mt + r) * 16 + 4 * q + j];
        w.w3[mt] = pack4(v[0], v[1], v[2], v[3]);
        for (int kb = 0; kb < 4; kb++) {
            for (int j = 0; j < 4; j++) v[j] = P[OFF_W2 + (16 * mt + r) * 64 + 16 * kb + 4 * q + j];
            w.w2[mt][kb] = pack4(v[0], v[1], v[2], v[3]);
        }
    }
}

// biases of features 16 mt + 4q + j as D-layout accumulators, held in registers
struct FwdB {
    f32x4 b1[4], b2[4], b3;
};

__device__ void load_bias(const float* __restrict__ P, FwdB& b, int lane) {
    const int q = lane / 16;
    for (int mt = 0; mt < 4; mt++)
        for (int j = 0; j < 4; j++) {
            b.b1[mt][j] = P[OFF_B1 + 16 * mt + 4 * q + j];
            b.b2[mt][j] = P[OFF_B2 + 16 * mt + 4 * q + j];
        }
    for (int j = 0; j < 4; j++) b.b3[j] = P[OFF_B3 + 4 * q + j];
}

// tanh x = sign(x) (1 - e) / (1 + e), e = exp(-2|x|) = 2^(-2 log2(e) |x|), on
// the hardware exp2 and reciprocal (about 1e-7 absolute error, far below the
// BF16 rounding that follows)
// tanh x = 1 - 2 / (1 + e^{2x}): five instructions (two of them
// transcendental), saturating to +-1 without a branch (e^{2x} overflows to
// +inf, whose reciprocal is 0); the absolute error is a few FP32 ulps of 1
__device__ __forceinline__ float tanh_fast(float x) {
    const float e = __builtin_amdgcn_exp2f(2.8853900817779268f * x);
    return fmaf(-2.f, __builtin_amdgcn_rcpf(1.f + e), 1.f);
}

__device__ __forceinline__ f32x4 tanh4(f32x4 v) {
    f32x4 r;
    for (int j = 0; j < 4; j++) r[j] = tanh_fast(v[j]);
    return r;
}

// Feature f of sample s.  Planar: float fields ld apart.  Packed: complex64
// fields ld elements apart, feature f is component f % 2 of field f / 2 (the
// physical values of two real fields transformed together).
template <bool PACKED>
__device__ __forceinline__ long fidx(int f, long s, long ld) {
    if constexpr (PACKED)
        return 2 * ((long)(f >> 1) * ld + s) + (f & 1);
    else
        return (long)f * ld + s;
}

// per-lane input scales of features 4q + j (zero for the padding features)
__device__ __forceinline__ f32x4 load_scale_in(const float* __restrict__ sc, int lane) {
    const int q = lane / 16;
    f32x4 r;
    for (int j = 0; j < 4; j++) r[j] = 4 * q + j < kIn ? sc[4 * q + j] : 0.f;
    return r;
}

// B operand X^T for samples s0..s0+15: lane holds features 4q+j of sample
// s0 + r, scaled; features >= 9 and samples >= n are zero
template <bool PACKED>
__device__ __forceinline__ s16x4 load_x(const float* __restrict__ X, long ld, long s0, long n, f32x4 sc,
                                        int lane, f32x4* raw = nullptr) {
    const long s = s0 + lane % 16;
    const int q = lane / 16;
    f32x4 xf;
    for (int j = 0; j < 4; j++) {
        const int f = 4 * q + j;
        xf[j] = (f < kIn && s < n) ? X[fidx<PACKED>(f, s, ld)] : 0.f;
    }
    if (raw) *raw = xf;  // unscaled inputs, for the clip test
    xf *= sc;
    return pack4(xf[0], xf[1], xf[2], xf[3]);
}

// the unscaled inputs of load_x (features 4q+j of sample s0 + r; zero for
// features >= 9 and samples >= n), without the conversion to an operand
template <bool PACKED>
__device__ __forceinline__ f32x4 load_raw(const float* __restrict__ X, long ld, long s0, long n, int lane) {
    const long s = s0 + lane % 16;
    const int q = lane / 16;
    f32x4 xf;
    for (int j = 0; j < 4; j++) {
        const int f = 4 * q + j;
        xf[j] = (f < kIn && s < n) ? X[fidx<PACKED>(f, s, ld)] : 0.f;
    }
    return xf;
}

struct Fwd {
    f32x4 h1[4], h2[4], y;
    s16x4 h1b[4], h2b[4];
};

__device__ __forceinline__ void forward(const FwdW& w, const FwdB& b, s16x4 xb, Fwd& o) {
    for (int mt = 0; mt < 4; mt++) {
        const f32x4 acc = mfma(w.w1t[mt], xb, b.b1[mt]);
        o.h1[mt] = tanh4(acc);
        o.h1b[mt] = to_bf(o.h1[mt]);
    }
    for (int mt = 0; mt < 4; mt++) {
        f32x4 acc = b.b2[mt];
        for (int kb = 0; kb < 4; kb++) acc = mfma(w.w2t[mt][kb], o.h1b[kb], acc);
        o.h2[mt] = tanh4(acc);
        o.h2b[mt] = to_bf(o.h2[mt]);
    }
    f32x4 acc = b.b3;
    for (int kb = 0; kb < 4; kb++) acc = mfma(w.w3t[kb], o.h2b[kb], acc);
    o.y = acc;
}

// two tiles through the network, interleaved layer by layer (independent
// chains for the scheduler)
__device__ __forceinline__ void forward2(const FwdW& w, const FwdB& b, s16x4 xa, s16x4 xb, f32x4& ya, f32x4& yb) {
    s16x4 ha[4], hb[4];
    for (int mt = 0; mt < 4; mt++) {
        const f32x4 a = mfma(w.w1t[mt], xa, b.b1[mt]);
        const f32x4 c = mfma(w.w1t[mt], xb, b.b1[mt]);
        ha[mt] = to_bf(tanh4(a));
        hb[mt] = to_bf(tanh4(c));
    }
    s16x4 ga[4], gb[4];
    for (int mt = 0; mt < 4; mt++) {
        f32x4 a = b.b2[mt], c = b.b2[mt];
        for (int kb = 0; kb < 4; kb++) {
            a = mfma(w.w2t[mt][kb], ha[kb], a);
            c = mfma(w.w2t[mt][kb], hb[kb], c);
        }
        ga[mt] = to_bf(tanh4(a));
        gb[mt] = to_bf(tanh4(c));
    }
    ya = b.b3;
    yb = b.b3;
    for (int kb = 0; kb < 4; kb++) {
        ya = mfma(w.w3t[kb], ga[kb], ya);
        yb = mfma(w.w3t[kb], gb[kb], yb);
    }
}

// Backscatter clipping of one D tile of stresses y (scaled outputs, lanes
// q = 0, 1 hold a sample's six) given the same tile's unscaled inputs raw
// (lanes q = 0, 1, 2 hold the nine gradients g_{3i+j}): zero where the
// deviatoric stress has tau_ij S_ij > 0, S the strain rate.  Every lane
// gathers its sample's values, so the whole wave must call it.
__device__ __forceinline__ f32x4 clip_backscatter(f32x4 y, f32x4 raw, int r) {
    float t[6];
    for (int j = 0; j < 4; j++) t[j] = __shfl(y[j], r, 64);
    t[4] = __shfl(y[0], r + 16, 64);
    t[5] = __shfl(y[1], r + 16, 64);
    float g[9];
    for (int f = 0; f < 9; f++) g[f] = __shfl(raw[f & 3], r + 16 * (f >> 2), 64);
    const float tr = (t[0] + t[1] + t[2]) * (1.f / 3.f);
    const float sxy = 0.5f * (g[1] + g[3]), sxz = 0.5f * (g[2] + g[6]), syz = 0.5f * (g[5] + g[7]);
    const float ts = (t[0] - tr) * g[0] + (t[1] - tr) * g[4] + (t[2] - tr) * g[8] +
                     2.f * (t[3] * sxy + t[4] * sxz + t[5] * syz);
    return ts > 0.f ? f32x4{0, 0, 0, 0} : y;
}

// Smagorinsky stress added to one D tile of stresses: y_o -= 2 c |S| S_o for
// the outputs o = 4q + j < 6 this lane holds (xx yy zz xy xz yz), c =
// (Cs Delta)^2, S the strain rate of the sample's nine unscaled gradients
// (gathered across the lanes as clip_backscatter does), |S| = sqrt(2 S:S).
// Every lane gathers, so the whole wave must call it.
__device__ __forceinline__ f32x4 smagorinsky_add(f32x4 y, f32x4 raw, int r, float c) {
    const int q = (threadIdx.x % 64) / 16;
    float g[9];
    for (int f = 0; f < 9; f++) g[f] = __shfl(raw[f & 3], r + 16 * (f >> 2), 64);
    const float s[6] = {g[0], g[4], g[8], 0.5f * (g[1] + g[3]), 0.5f * (g[2] + g[6]), 0.5f * (g[5] + g[7])};
    const float ss = s[0] * s[0] + s[1] * s[1] + s[2] * s[2] + 2.f * (s[3] * s[3] + s[4] * s[4] + s[5] * s[5]);
    const float nut = c * sqrtf(2.f * ss);
    for (int j = 0; j < 4; j++) {
        const float so = q == 0 ? s[j] : (q == 1 && j < 2) ? s[4 + j] : 0.f;
        y[j] -= 2.f * nut * so;
    }
    return y;
}

}  // namespace sh
