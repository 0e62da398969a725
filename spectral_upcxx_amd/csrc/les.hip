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
//   4  forward 3D transform of the stresses (3 fields, not 6)
//   5  unpack by Hermitian symmetry, f_i = -sum_j i k_j tau_ij
// Nyquist wavenumbers count as zero in the derivatives, so every packed field
// is the transform of real fields and the packing is exact.
#include <cmath>

#include "sh_common.hpp"

namespace sh {

template <typename T>
int fft3d(void* data, int count, long fstride, int nx, int ny, int nz, int dir, int flags, void* stream);
int mlp_infer_packed(sh_mlp_t h, const float2* X, long ld, long n, float2* Y, long ldy, int clip, hipStream_t st);
int mlp_infer_packed_planar(sh_mlp_t h, const float2* X, long ld, long n, float* Y, long ldy, int clip,
                            hipStream_t st);
void* ns_fused_velocity(sh_ns_t s);

// wavenumber of index i on an n-point axis, with the Nyquist index as zero
__device__ __forceinline__ float kd(int i, int n) { return (2 * i == n) ? 0.f : (float)(i <= n / 2 ? i : i - n); }

__device__ __forceinline__ float2 ik(float k, float2 v) { return make_float2(-k * v.y, k * v.x); }

// Z_m = g_{2m} + i g_{2m+1} with g_{3a+b} = i k_b V_a
__global__ void k_les_grad_pack(const float2* __restrict__ V, long fs, float2* __restrict__ Z, long zs, int n) {
    const long N = (long)n * n * n;
    for (long p = blockIdx.x * (long)blockDim.x + threadIdx.x; p < N; p += (long)gridDim.x * blockDim.x) {
        const int i = (int)(p % n), j = (int)((p / n) % n), k = (int)(p / ((long)n * n));
        const float kk[3] = {kd(i, n), kd(j, n), kd(k, n)};
        float2 g[10];
        for (int a = 0; a < 3; a++) {
            const float2 v = V[a * fs + p];
            for (int b = 0; b < 3; b++) g[3 * a + b] = ik(kk[b], v);
        }
        g[9] = make_float2(0.f, 0.f);
        for (int m = 0; m < 5; m++) {
            const float2 a = g[2 * m], b = g[2 * m + 1];
            Z[m * zs + p] = make_float2(a.x - b.y, a.y + b.x);
        }
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

// physical packed gradient fields Z (5 complex64 fields) of V (3 Fourier fields)
int les_gradients(const float2* V, float2* Z, int n, hipStream_t st) {
    const long N = (long)n * n * n;
    hipLaunchKernelGGL(k_les_grad_pack, dim3(lgrid(N)), dim3(256), 0, st, V, N, Z, N, n);
    SH_LAUNCH_CHECK("k_les_grad_pack");
    {
        ProfScope ps("les_fft_inv", st);
        RC(fft3d<float>(Z, 5, N, n, n, n, +1, SH_NORMALIZE, st));
    }
    return SH_OK;
}

}  // namespace sh

using namespace sh;

struct sh_les_s {
    int n;
    long N;
    sh_mlp_t model;
    int clip;
    hipStream_t st;
    sh_ns_t ns;
    float2 *F, *Z, *T, *Vc;  // force (3), packed gradients (5), packed stresses (3), FP32 velocity (3)
    double2* V64;            // 3 fields for the initial condition
};

namespace {

int sgs_force(sh_les_s* h, const float2* V) {
    const long N = h->N;
    RC(les_gradients(V, h->Z, h->n, h->st));
    {
        ProfScope ps("les_mlp", h->st);
        RC(mlp_infer_packed(h->model, h->Z, N, N, h->T, N, h->clip, h->st));
    }
    {
        ProfScope ps("les_fft_fwd", h->st);
        RC(fft3d<float>(h->T, 3, N, h->n, h->n, h->n, -1, 0, h->st));
    }
    hipLaunchKernelGGL(k_les_div_unpack, dim3(lgrid(N)), dim3(256), 0, h->st, h->T, N, h->F, N, h->n);
    SH_LAUNCH_CHECK("k_les_div_unpack");
    return SH_OK;
}

void free_les(sh_les_s* h) {
    if (h->ns) sh_ns_destroy(h->ns);
    float2* bufs[4] = {h->F, h->Z, h->T, h->Vc};
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
    int rc = sh_ns_create(&h->ns, n, n, n, nu, K, SH_F32, 1, stream);
    void* p;
    const long sizes[4] = {3, 5, 3, 3};
    float2** dst[4] = {&h->F, &h->Z, &h->T, &h->Vc};
    for (int i = 0; i < 4; i++) {
        p = nullptr;
        if (!rc) rc = pool_alloc(&p, sizeof(float2) * h->N * sizes[i]);
        *dst[i] = (float2*)p;
    }
    p = nullptr;
    if (!rc) rc = pool_alloc(&p, sizeof(double2) * h->N * 3);
    h->V64 = (double2*)p;
    if (!rc && model) rc = sh_ns_set_forcing(h->ns, h->F);
    if (rc) {
        free_les(h);
        return rc;
    }
    *handle = h;
    return SH_OK;
}

int sh_les_destroy(sh_les_t h) {
    if (!h) return SH_OK;
    (void)hipStreamSynchronize(h->st);
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
        RC(sgs_force(h, h->Vc));
    }
    return sh_ns_startup(h->ns);
}

int sh_les_steps(sh_les_t h, int nsteps) {
    if (!h || nsteps < 0) return SH_EINVAL;
    const float2* V = (const float2*)ns_fused_velocity(h->ns);
    for (int t = 0; t < nsteps; t++) {
        if (h->model) RC(sgs_force(h, V));
        RC(sh_ns_steps(h->ns, 1));
    }
    return SH_OK;
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
    RC(les_gradients((const float2*)ns_fused_velocity(h->ns), h->Z, h->n, h->st));
    return mlp_infer_packed_planar(h->model, h->Z, h->N, h->N, (float*)out_tau6, h->N, h->clip, h->st);
}

int sh_les_force(sh_les_t h, void* out_f3) {
    if (!h || !out_f3 || !h->model) {
        set_error("sh_les_force: invalid argument (an LES with a model)");
        return SH_EINVAL;
    }
    RC(sgs_force(h, (const float2*)ns_fused_velocity(h->ns)));
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
    if (!rc) rc = les_gradients((const float2*)v, (float2*)z, n, st);
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
