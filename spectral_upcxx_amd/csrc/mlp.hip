// mlp.hip -- the learned subgrid-stress closure: a 9 -> 64 -> 64 -> 6 MLP
// (tanh hidden layers) on BF16 MFMA with FP32 accumulation, and its training
// (forward, backward, Adam on FP32 master weights).  No counterpart in the
// reference.  The operand layout and the shared device pieces are in
// mlp_core.hpp; the weight gradients reduce over samples, their operands
// come from 16x16 transposes of D tiles through a small LDS tile.
#include <cmath>
#include <vector>

#include "mlp_core.hpp"

namespace sh {

// Inference over samples [0, n): Y = out_scale * net(in_scale * X).  With
// clip, a sample's stress is zeroed where its deviatoric part has
// tau_ij S_ij > 0 (S the strain rate of the gradient inputs g_{3i+j}).
template <bool PACKED_IN, bool PACKED_OUT>
__global__ __launch_bounds__(256) void k_mlp_infer(const float* __restrict__ P, const float* __restrict__ X,
                                                   long ld, long n, const float* __restrict__ sc,
                                                   float* __restrict__ Y, long ldy, int clip, float smag) {
    const int lane = threadIdx.x % 64;
    const int r = lane % 16, q = lane / 16;
    const long wave = (blockIdx.x * (long)blockDim.x + threadIdx.x) / 64;
    const long nwaves = (long)gridDim.x * blockDim.x / 64;
    FwdW w;
    FwdB bias;
    load_fwd(P, w, lane);
    load_bias(P, bias, lane);
    const f32x4 sc_in = load_scale_in(sc, lane);
    f32x4 sc_out;
    for (int j = 0; j < 4; j++) sc_out[j] = 4 * q + j < kOut ? sc[kIn + 4 * q + j] : 0.f;
    // two tiles per trip: their independent MFMA -> tanh -> convert chains
    // interleave, which hides the latency one tile alone leaves exposed; the
    // next trip's inputs are loaded before this trip's network runs
    f32x4 na = load_raw<PACKED_IN>(X, ld, wave * 32, n, lane);
    f32x4 nb = load_raw<PACKED_IN>(X, ld, wave * 32 + 16, n, lane);
    for (long s00 = wave * 32; s00 < n; s00 += nwaves * 32) {
        const f32x4 ra = na, rb = nb;
        na = load_raw<PACKED_IN>(X, ld, s00 + nwaves * 32, n, lane);
        nb = load_raw<PACKED_IN>(X, ld, s00 + nwaves * 32 + 16, n, lane);
        f32x4 ya, yb;
        forward2(w, bias, to_bf(ra * sc_in), to_bf(rb * sc_in), ya, yb);
        for (int half = 0; half < 2; half++) {
            const long s = s00 + 16 * half + r;
            f32x4 y = (half ? yb : ya) * sc_out;
            if (clip) y = clip_backscatter(y, half ? rb : ra, r);
            if (smag > 0.f) y = smagorinsky_add(y, half ? rb : ra, r, smag);
            if (s < n && q < 2) {
                if constexpr (PACKED_OUT) {
                    float2* Y2 = reinterpret_cast<float2*>(Y);
                    Y2[(2 * q) * ldy + s] = make_float2(y[0], y[1]);
                    if (q == 0) Y2[ldy + s] = make_float2(y[2], y[3]);
                } else {
                    for (int j = 0; j < 4; j++)
                        if (4 * q + j < kOut) Y[(4 * q + j) * ldy + s] = y[j];
                }
            }
        }
    }
}

// 16x16 transpose of a D-layout tile through the block's LDS tile (the block
// is one wavefront): returns the A/B operand Q[r][4q + j] of the tile Q whose
// D layout is d.
__device__ __forceinline__ s16x4 transpose(s16x4 d, short* t, int lane) {
    const int r = lane % 16, q = lane / 16;
    for (int j = 0; j < 4; j++) t[(4 * q + j) * 17 + r] = d[j];
    __syncthreads();
    s16x4 o;
    for (int j = 0; j < 4; j++) o[j] = t[r * 17 + 4 * q + j];
    __syncthreads();
    return o;
}

// Gradient of the mean-square loss over samples [0, n): each wave (one per
// block) accumulates its tiles' gradient and squared error into
// part[wave][NPARAM + 1].  Targets are divided by the output scale (the
// network predicts stress / out_scale).
// DY_IN: the loss gradient with respect to the network outputs is given
// (Tg holds dL/dy, 6 planar fields ld apart) instead of being formed from
// targets; the squared-error slot is then left zero.
template <bool DY_IN>
__global__ __launch_bounds__(64) void k_mlp_grad(const float* __restrict__ P, const float* __restrict__ X,
                                                 const float* __restrict__ Tg, long ld, long n,
                                                 const float* __restrict__ sc, float inv_count,
                                                 float* __restrict__ part) {
    __shared__ short tile[16 * 17];
    const int lane = threadIdx.x;
    const int r = lane % 16, q = lane / 16;
    const long wave = blockIdx.x;
    const long nwaves = gridDim.x;
    FwdW w;
    BwdW wb;
    FwdB bias;
    load_fwd(P, w, lane);
    load_bwd(P, wb, lane);
    load_bias(P, bias, lane);
    const f32x4 sc_in = load_scale_in(sc, lane);
    f32x4 inv_out;  // targets are divided by the output scale
    for (int j = 0; j < 4; j++) inv_out[j] = 4 * q + j < kOut ? 1.f / sc[kIn + 4 * q + j] : 0.f;
    f32x4 gw1[4], gw2[4][4], gw3[4], gb1[4], gb2[4], gb3;
    for (int a = 0; a < 4; a++) {
        gw1[a] = gw3[a] = gb1[a] = gb2[a] = f32x4{0, 0, 0, 0};
        for (int b = 0; b < 4; b++) gw2[a][b] = f32x4{0, 0, 0, 0};
    }
    gb3 = f32x4{0, 0, 0, 0};
    float sq = 0;
    for (long s0 = wave * 16; s0 < n; s0 += nwaves * 16) {
        const s16x4 xb = load_x<false>(X, ld, s0, n, sc_in, lane);
        Fwd o;
        forward(w, bias, xb, o);
        // dY = 2 (Y - T) / count on valid outputs and samples
        const long s = s0 + r;
        f32x4 dy;
        for (int j = 0; j < 4; j++) {
            const int f = 4 * q + j;
            float d = 0;
            if (f < kOut && s < n) {
                if constexpr (DY_IN) {
                    d = Tg[f * ld + s];
                } else {
                    const float e = o.y[j] - Tg[f * ld + s] * inv_out[j];
                    sq += e * e;
                    d = 2.f * e * inv_count;
                }
            }
            dy[j] = d;
        }
        const s16x4 dyb = to_bf(dy);
        gb3 += dy;
        // dZ2 = (W3 dY) (1 - H2^2)
        s16x4 dz2b[4];
        for (int mt = 0; mt < 4; mt++) {
            f32x4 acc = mfma(wb.w3[mt], dyb, f32x4{0, 0, 0, 0});
            for (int j = 0; j < 4; j++) acc[j] *= 1.f - o.h2[mt][j] * o.h2[mt][j];
            dz2b[mt] = to_bf(acc);
            gb2[mt] += acc;
        }
        // dZ1 = (W2 dZ2) (1 - H1^2)
        s16x4 dz1b[4];
        for (int mt = 0; mt < 4; mt++) {
            f32x4 acc = f32x4{0, 0, 0, 0};
            for (int kb = 0; kb < 4; kb++) acc = mfma(wb.w2[mt][kb], dz2b[kb], acc);
            for (int j = 0; j < 4; j++) acc[j] *= 1.f - o.h1[mt][j] * o.h1[mt][j];
            dz1b[mt] = to_bf(acc);
            gb1[mt] += acc;
        }
        // dW[a][b] = sum_s act_s[a] delta_s[b]: K runs over the 16 samples
        const s16x4 xT = transpose(xb, tile, lane);
        const s16x4 dyT = transpose(dyb, tile, lane);
        s16x4 h1T[4];
        for (int t = 0; t < 4; t++) {
            const s16x4 dz1T = transpose(dz1b[t], tile, lane);
            gw1[t] = mfma(xT, dz1T, gw1[t]);
            const s16x4 h2T = transpose(o.h2b[t], tile, lane);
            gw3[t] = mfma(h2T, dyT, gw3[t]);
            h1T[t] = transpose(o.h1b[t], tile, lane);
        }
        for (int bt = 0; bt < 4; bt++) {
            const s16x4 dz2T = transpose(dz2b[bt], tile, lane);
            for (int at = 0; at < 4; at++) gw2[at][bt] = mfma(h1T[at], dz2T, gw2[at][bt]);
        }
    }
    // D layout of gradient tile (at, bt): lane holds dW[16 at + 4q + j][16 bt + r]
    float* out = part + wave * (NPARAM + 1);
    for (int bt = 0; bt < 4; bt++)
        for (int j = 0; j < 4; j++)
            out[OFF_W1 + (4 * q + j) * 64 + 16 * bt + r] = (4 * q + j < kIn) ? gw1[bt][j] : 0.f;
    for (int at = 0; at < 4; at++)
        for (int bt = 0; bt < 4; bt++)
            for (int j = 0; j < 4; j++) out[OFF_W2 + (16 * at + 4 * q + j) * 64 + 16 * bt + r] = gw2[at][bt][j];
    for (int at = 0; at < 4; at++)
        for (int j = 0; j < 4; j++) out[OFF_W3 + (16 * at + 4 * q + j) * 16 + r] = (r < kOut) ? gw3[at][j] : 0.f;
    // bias gradients: sums over the 16 samples, i.e. over the lanes of one q
    for (int mt = 0; mt < 4; mt++)
        for (int j = 0; j < 4; j++) {
            float v1 = gb1[mt][j], v2 = gb2[mt][j];
            for (int m = 1; m < 16; m <<= 1) {
                v1 += __shfl_xor(v1, m, 64);
                v2 += __shfl_xor(v2, m, 64);
            }
            if (r == 0) {
                out[OFF_B1 + 16 * mt + 4 * q + j] = v1;
                out[OFF_B2 + 16 * mt + 4 * q + j] = v2;
            }
        }
    for (int j = 0; j < 4; j++) {
        float v = gb3[j];
        for (int m = 1; m < 16; m <<= 1) v += __shfl_xor(v, m, 64);
        if (r == 0) out[OFF_B3 + 4 * q + j] = (4 * q + j < kOut) ? v : 0.f;
    }
    for (int m = 1; m < 64; m <<= 1) sq += __shfl_xor(sq, m, 64);
    if (lane == 0) out[NPARAM] = sq;
}

// sum of the per-wave partials in a fixed order (deterministic)
__global__ void k_mlp_reduce(const float* __restrict__ part, int nwaves, float* __restrict__ g) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i > NPARAM) return;
    double acc = 0;
    for (int w = 0; w < nwaves; w++) acc += part[(long)w * (NPARAM + 1) + i];
    g[i] = (float)acc;
}

// Adam on the FP32 master parameters; padding entries stay zero
__global__ void k_mlp_adam(float* __restrict__ P, float* __restrict__ m, float* __restrict__ v,
                           const float* __restrict__ g, float lr, float b1, float b2, float eps,
                           float c1, float c2) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= NPARAM) return;
    bool pad = false;
    if (i < OFF_B1)
        pad = (i / 64) >= kIn;
    else if (i >= OFF_W3 && i < OFF_B3)
        pad = ((i - OFF_W3) % 16) >= kOut;
    else if (i >= OFF_B3)
        pad = (i - OFF_B3) >= kOut;
    if (pad) {
        P[i] = 0;
        return;
    }
    const float gi = g[i];
    const float mi = b1 * m[i] + (1 - b1) * gi;
    const float vi = b2 * v[i] + (1 - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    P[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// D = A B through one MFMA with the operand layout stated at the top
__global__ void k_mfma_probe(const float* __restrict__ A, const float* __restrict__ B, float* __restrict__ D) {
    const int lane = threadIdx.x;
    const int r = lane % 16, q = lane / 16;
    const float* a = A + r * 16 + 4 * q;
    const float* b = B + 4 * q * 16 + r;
    const f32x4 d = mfma(pack4(a[0], a[1], a[2], a[3]), pack4(b[0], b[16], b[32], b[48]), f32x4{0, 0, 0, 0});
    for (int j = 0; j < 4; j++) D[(4 * q + j) * 16 + r] = d[j];
}

// at least 8 tiles per wave (the weights are loaded once per wave), 4 waves
// per block, and at most 2 waves per SIMD of the 256 CUs, all resident at
// once (the inference kernel needs about 150 VGPRs, so 3 would fit)
static unsigned infer_blocks(long n) {
    const long tiles = (n + 15) / 16;
    long b = (tiles + 8 * 4 - 1) / (8 * 4);
    if (b > 512) b = 512;
    return (unsigned)(b < 1 ? 1 : b);
}

}  // namespace sh

using namespace sh;

struct sh_mlp_s {
    float* P;      // NPARAM master weights
    float* m;      // Adam moments
    float* v;
    float* grad;   // NPARAM + 1 (the last: sum of squared errors)
    float* part;   // per-wave partials
    float* scale;  // [9] input scale, [6] output scale
    long t;        // Adam step count
    hipStream_t st;
    void* aux;          // scratch of the force-matching loss (grown on demand)
    size_t aux_bytes;
};

namespace sh {
// the LES feeds packed complex gradient fields and takes packed stresses
int mlp_infer_packed(sh_mlp_t h, const float2* X, long ld, long n, float2* Y, long ldy, int clip, float smag,
                     hipStream_t st) {
    hipLaunchKernelGGL((k_mlp_infer<true, true>), dim3(infer_blocks(n)), dim3(256), 0, st, h->P,
                       (const float*)X, ld, n, h->scale, (float*)Y, ldy, clip, smag);
    SH_LAUNCH_CHECK("k_mlp_infer");
    return SH_OK;
}
int mlp_infer_packed_planar(sh_mlp_t h, const float2* X, long ld, long n, float* Y, long ldy, int clip, float smag,
                            hipStream_t st) {
    hipLaunchKernelGGL((k_mlp_infer<true, false>), dim3(infer_blocks(n)), dim3(256), 0, st, h->P,
                       (const float*)X, ld, n, h->scale, Y, ldy, clip, smag);
    SH_LAUNCH_CHECK("k_mlp_infer");
    return SH_OK;
}
// planar features in, packed stresses out (the force-matching loss)
int mlp_infer_planar_packed(sh_mlp_t h, const float* X, long ld, long n, float2* Y, long ldy, hipStream_t st) {
    hipLaunchKernelGGL((k_mlp_infer<false, true>), dim3(infer_blocks(n)), dim3(256), 0, st, h->P, X, ld, n,
                       h->scale, (float*)Y, ldy, 0, 0.f);
    SH_LAUNCH_CHECK("k_mlp_infer");
    return SH_OK;
}
// parameter gradient for a given dL/dy (6 planar fields ld apart) into h->grad
int mlp_grad_dy(sh_mlp_t h, const float* X, const float* Dy, long ld, long n, hipStream_t st) {
    const long tiles = (n + 15) / 16;
    const int nw = (int)(tiles < kGradWaves ? tiles : kGradWaves);
    hipLaunchKernelGGL(k_mlp_grad<true>, dim3(nw), dim3(64), 0, st, h->P, X, Dy, ld, n, h->scale, 0.f, h->part);
    SH_LAUNCH_CHECK("k_mlp_grad");
    hipLaunchKernelGGL(k_mlp_reduce, dim3((NPARAM + 256) / 256), dim3(256), 0, st, h->part, nw, h->grad);
    SH_LAUNCH_CHECK("k_mlp_reduce");
    return SH_OK;
}
// one Adam step on h->grad
int mlp_adam(sh_mlp_t h, float lr, hipStream_t st) {
    h->t++;
    const double b1 = 0.9, b2 = 0.999;
    const float c1 = (float)(1.0 - std::pow(b1, (double)h->t)), c2 = (float)(1.0 - std::pow(b2, (double)h->t));
    hipLaunchKernelGGL(k_mlp_adam, dim3((NPARAM + 255) / 256), dim3(256), 0, st, h->P, h->m, h->v, h->grad, lr,
                       (float)b1, (float)b2, 1e-8f, c1, c2);
    SH_LAUNCH_CHECK("k_mlp_adam");
    return SH_OK;
}
// scratch of at least `bytes` held by the handle (previous contents dropped)
void* mlp_aux(sh_mlp_t h, size_t bytes) {
    if (h->aux_bytes < bytes) {
        if (h->aux) {
            if (hipStreamSynchronize(h->st) != hipSuccess) return nullptr;
            pool_free(h->aux);
            h->aux = nullptr;
            h->aux_bytes = 0;
        }
        if (pool_alloc(&h->aux, bytes)) return nullptr;
        h->aux_bytes = bytes;
    }
    return h->aux;
}
float* mlp_grad_buffer(sh_mlp_t h) { return h->grad; }
const float* mlp_scales(sh_mlp_t h) { return h->scale; }
const float* mlp_params(sh_mlp_t h) { return h->P; }
hipStream_t mlp_stream(sh_mlp_t h) { return h->st; }
}  // namespace sh

extern "C" {

int sh_mlp_param_count(int* count) {
    if (!count) return SH_EINVAL;
    *count = NPARAM;
    return SH_OK;
}

int sh_mlp_create(sh_mlp_t* handle, uint64_t seed, void* stream) {
    if (!handle) return SH_EINVAL;
    int rc = ensure_init();
    if (rc) return rc;
    sh_mlp_s* h = new sh_mlp_s();
    h->st = as_stream(stream);
    h->t = 0;
    const size_t need[6] = {NPARAM, NPARAM, NPARAM, NPARAM + 1, (size_t)kGradWaves * (NPARAM + 1), 15};
    float** dst[6] = {&h->P, &h->m, &h->v, &h->grad, &h->part, &h->scale};
    for (int i = 0; i < 6; i++) {
        void* p = nullptr;
        if (!rc) rc = pool_alloc(&p, need[i] * sizeof(float));
        *dst[i] = (float*)p;
    }
    if (rc) {
        for (int i = 0; i < 6; i++)
            if (*dst[i]) pool_free(*dst[i]);
        delete h;
        return rc;
    }
    // Glorot-uniform weights from the splitmix64 stream, zero biases and padding
    std::vector<float> P(NPARAM, 0.f);
    uint64_t c = 0;
    auto u01 = [&]() {
        uint64_t z = seed + (++c) * 0x9E3779B97F4A7C15ULL;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
        z ^= z >> 31;
        return (double)(z >> 11) * 0x1.0p-53;
    };
    auto fill = [&](int off, int cols, int vr, int vc) {
        const double a = std::sqrt(6.0 / (vr + vc));
        for (int i = 0; i < vr; i++)
            for (int j = 0; j < vc; j++) P[off + i * cols + j] = (float)((2 * u01() - 1) * a);
    };
    fill(OFF_W1, 64, kIn, kH);
    fill(OFF_W2, 64, kH, kH);
    fill(OFF_W3, 16, kH, kOut);
    std::vector<float> ones(15, 1.f);
    SH_HIP(hipMemcpy(h->P, P.data(), NPARAM * sizeof(float), hipMemcpyHostToDevice));
    SH_HIP(hipMemset(h->m, 0, NPARAM * sizeof(float)));
    SH_HIP(hipMemset(h->v, 0, NPARAM * sizeof(float)));
    SH_HIP(hipMemcpy(h->scale, ones.data(), 15 * sizeof(float), hipMemcpyHostToDevice));
    *handle = h;
    return SH_OK;
}

int sh_mlp_destroy(sh_mlp_t h) {
    if (!h) return SH_OK;
    (void)hipStreamSynchronize(h->st);
    pool_free(h->P);
    pool_free(h->m);
    pool_free(h->v);
    pool_free(h->grad);
    pool_free(h->part);
    pool_free(h->scale);
    if (h->aux) pool_free(h->aux);
    delete h;
    return SH_OK;
}

int sh_mlp_set_params(sh_mlp_t h, const float* host_params) {
    if (!h || !host_params) return SH_EINVAL;
    SH_HIP(hipStreamSynchronize(h->st));
    SH_HIP(hipMemcpy(h->P, host_params, NPARAM * sizeof(float), hipMemcpyHostToDevice));
    return SH_OK;
}

int sh_mlp_get_params(sh_mlp_t h, float* host_params) {
    if (!h || !host_params) return SH_EINVAL;
    SH_HIP(hipStreamSynchronize(h->st));
    SH_HIP(hipMemcpy(host_params, h->P, NPARAM * sizeof(float), hipMemcpyDeviceToHost));
    return SH_OK;
}

int sh_mlp_set_scaling(sh_mlp_t h, const float* in_scale9, const float* out_scale6) {
    if (!h || !in_scale9 || !out_scale6) return SH_EINVAL;
    SH_HIP(hipStreamSynchronize(h->st));
    SH_HIP(hipMemcpy(h->scale, in_scale9, 9 * sizeof(float), hipMemcpyHostToDevice));
    SH_HIP(hipMemcpy(h->scale + 9, out_scale6, 6 * sizeof(float), hipMemcpyHostToDevice));
    return SH_OK;
}

int sh_mlp_forward(sh_mlp_t h, const void* X, long ld, long n, void* Y, long ldy) {
    if (!h || !X || !Y || n < 0 || ld < n || ldy < n) {
        set_error("sh_mlp_forward: invalid argument");
        return SH_EINVAL;
    }
    if (n == 0) return SH_OK;
    hipLaunchKernelGGL((k_mlp_infer<false, false>), dim3(infer_blocks(n)), dim3(256), 0, h->st, h->P,
                       (const float*)X, ld, n, h->scale, (float*)Y, ldy, 0, 0.f);
    SH_LAUNCH_CHECK("k_mlp_infer");
    return SH_OK;
}

int sh_mlp_gradient(sh_mlp_t h, const void* X, const void* T, long ld, long n, void* grad_out, double* loss) {
    if (!h || !X || !T || n <= 0 || ld < n) {
        set_error("sh_mlp_gradient: invalid argument");
        return SH_EINVAL;
    }
    const long tiles = (n + 15) / 16;
    const int nw = (int)(tiles < kGradWaves ? tiles : kGradWaves);
    hipLaunchKernelGGL(k_mlp_grad<false>, dim3(nw), dim3(64), 0, h->st, h->P, (const float*)X, (const float*)T, ld,
                       n, h->scale, (float)(1.0 / ((double)n * kOut)), h->part);
    SH_LAUNCH_CHECK("k_mlp_grad");
    hipLaunchKernelGGL(k_mlp_reduce, dim3((NPARAM + 256) / 256), dim3(256), 0, h->st, h->part, nw, h->grad);
    SH_LAUNCH_CHECK("k_mlp_reduce");
    if (grad_out)
        SH_HIP(hipMemcpyAsync(grad_out, h->grad, NPARAM * sizeof(float), hipMemcpyDeviceToDevice, h->st));
    if (loss) {
        float sq = 0;
        SH_HIP(hipMemcpyAsync(&sq, h->grad + NPARAM, sizeof(float), hipMemcpyDeviceToHost, h->st));
        SH_HIP(hipStreamSynchronize(h->st));
        *loss = sq / ((double)n * kOut);
    }
    return SH_OK;
}

int sh_mlp_train_step(sh_mlp_t h, const void* X, const void* T, long ld, long n, float lr, double* loss) {
    int rc = sh_mlp_gradient(h, X, T, ld, n, nullptr, loss);
    if (rc) return rc;
    return mlp_adam(h, lr, h->st);
}

int sh_mfma_probe(const void* A, const void* B, void* D, void* stream) {
    if (!A || !B || !D) return SH_EINVAL;
    int rc = ensure_init();
    if (rc) return rc;
    hipLaunchKernelGGL(k_mfma_probe, dim3(1), dim3(64), 0, as_stream(stream), (const float*)A, (const float*)B,
                       (float*)D);
    SH_LAUNCH_CHECK("k_mfma_probe");
    return SH_OK;
}

}  // extern "C"
