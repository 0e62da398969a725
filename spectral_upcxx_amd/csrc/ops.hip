// ops.hip -- spectral operators (src/Scalar.cpp:245-370, src/Vector.cpp:66-181),
// elementwise Operators (src/Operators.cpp:9-99), reductions
// (src/Scalar.cpp:15-22), initial conditions and spectral diagnostics.
//
// All of these are HBM-streaming kernels (<= 0.5 flop/B): one complex element
// per thread per grid-stride trip, 16-byte loads, wavenumbers recomputed from
// the linear index (Scalar::freq) instead of being read from memory.
#include <cmath>
#include <mutex>
#include <string>
#include <vector>

#include "sh_common.hpp"

namespace sh {

template <typename T>
int fft3d(void* data, int count, long fstride, int nx, int ny, int nz, int dir, int flags,
          void* stream);

static constexpr int kBlock = 256;

static unsigned grid_for(long n) {
    long g = (n + kBlock - 1) / kBlock;
    if (g > 65536L * 4) g = 65536L * 4;
    if (g < 1) g = 1;
    return (unsigned)g;
}

struct Dims {
    int nx, ny, nz;
    __host__ __device__ long n() const { return (long)nx * ny * nz; }
};

template <typename T>
__device__ __forceinline__ void wavenumber(long n, const Dims& d, T& k0, T& k1, T& k2) {
    const int i = (int)(n % d.nx);
    const long r = n / d.nx;
    const int j = (int)(r % d.ny);
    const int k = (int)(r / d.ny);
    k0 = freq<T>(i, d.nx);
    k1 = freq<T>(j, d.ny);
    k2 = freq<T>(k, d.nz);
}

#define GRID_STRIDE(n, N) for (long n = blockIdx.x * (long)blockDim.x + threadIdx.x; n < (N); n += (long)gridDim.x * blockDim.x)

// ---------------------------------------------------------------- spectral operators
__global__ void k_gradient(const double2* __restrict__ f, double2* __restrict__ vx,
                           double2* __restrict__ vy, double2* __restrict__ vz, Dims d) {
    GRID_STRIDE(n, d.n()) {
        double k0, k1, k2;
        wavenumber(n, d, k0, k1, k2);
        const double2 a = f[n];
        vx[n] = imul(k0, a);
        vy[n] = imul(k1, a);
        vz[n] = imul(k2, a);
    }
}

__global__ void k_laplacian(const double2* __restrict__ f, double2* __restrict__ out, Dims d) {
    GRID_STRIDE(n, d.n()) {
        double k0, k1, k2;
        wavenumber(n, d, k0, k1, k2);
        const double q = -(k0 * k0 + k1 * k1 + k2 * k2);
        out[n] = rmul(q, f[n]);
    }
}

__global__ void k_laplacian_inverse(double2* __restrict__ f, double a, double b, Dims d) {
    GRID_STRIDE(n, d.n()) {
        double k0, k1, k2;
        wavenumber(n, d, k0, k1, k2);
        double sum = 0;
        sum += k0 * k0;
        sum += k1 * k1;
        sum += k2 * k2;
        if (sum != 0) {
            const double den = a - b * sum;
            double2 v = f[n];
            v.x /= den;
            v.y /= den;
            f[n] = v;
        }
    }
}

__global__ void k_divergence(const double2* __restrict__ vx, const double2* __restrict__ vy,
                             const double2* __restrict__ vz, double2* __restrict__ out, Dims d) {
    GRID_STRIDE(n, d.n()) {
        double k0, k1, k2;
        wavenumber(n, d, k0, k1, k2);
        double2 acc = mk<double>(0, 0);
        acc = cadd(acc, imul(k0, vx[n]));
        acc = cadd(acc, imul(k1, vy[n]));
        acc = cadd(acc, imul(k2, vz[n]));
        out[n] = acc;
    }
}

__global__ void k_curl(const double2* __restrict__ vx, const double2* __restrict__ vy,
                       const double2* __restrict__ vz, double2* __restrict__ ox,
                       double2* __restrict__ oy, double2* __restrict__ oz, Dims d) {
    GRID_STRIDE(n, d.n()) {
        double k0, k1, k2;
        wavenumber(n, d, k0, k1, k2);
        const double2 a = vx[n], b = vy[n], c = vz[n];
        ox[n] = csub(imul(k1, c), imul(k2, b));
        oy[n] = csub(imul(k2, a), imul(k0, c));
        oz[n] = csub(imul(k0, b), imul(k1, a));
    }
}

__global__ void k_vlaplacian(const double2* __restrict__ vx, const double2* __restrict__ vy,
                             const double2* __restrict__ vz, double2* __restrict__ ox,
                             double2* __restrict__ oy, double2* __restrict__ oz, Dims d) {
    GRID_STRIDE(n, d.n()) {
        double k0, k1, k2;
        wavenumber(n, d, k0, k1, k2);
        const double q = -(k0 * k0 + k1 * k1 + k2 * k2);
        ox[n] = rmul(q, vx[n]);
        oy[n] = rmul(q, vy[n]);
        oz[n] = rmul(q, vz[n]);
    }
}

__global__ void k_zero_high(double2* __restrict__ f, Dims d) {
    GRID_STRIDE(n, d.n()) {
        const int i = (int)(n % d.nx);
        const long r = n / d.nx;
        const int j = (int)(r % d.ny);
        const int k = (int)(r / d.ny);
        if (i == d.nx / 2 || j == d.ny / 2 || k == d.nz / 2) f[n] = mk<double>(0, 0);
    }
}

// ret[0] = u1 v2 - u2 v1, ret[1] = u2 v0 - u0 v2, ret[2] = u0 v1 - u1 v0 (complex)
__global__ void k_cross_pw(const double2* __restrict__ ux, const double2* __restrict__ uy,
                           const double2* __restrict__ uz, const double2* __restrict__ wx,
                           const double2* __restrict__ wy, const double2* __restrict__ wz,
                           double2* __restrict__ ox, double2* __restrict__ oy,
                           double2* __restrict__ oz, long N) {
    GRID_STRIDE(n, N) {
        const double2 u0 = ux[n], u1 = uy[n], u2 = uz[n];
        const double2 v0 = wx[n], v1 = wy[n], v2 = wz[n];
        ox[n] = csub(cmul(u1, v2), cmul(u2, v1));
        oy[n] = csub(cmul(u2, v0), cmul(u0, v2));
        oz[n] = csub(cmul(u0, v1), cmul(u1, v0));
    }
}

// ---------------------------------------------------------------- elementwise
__global__ void k_mul(double2* __restrict__ o, const double2* __restrict__ x,
                      const double2* __restrict__ y, long N) {
    GRID_STRIDE(n, N) o[n] = cmul(x[n], y[n]);
}
__global__ void k_add(double2* __restrict__ o, const double2* __restrict__ x,
                      const double2* __restrict__ y, long N) {
    GRID_STRIDE(n, N) o[n] = cadd(x[n], y[n]);
}
__global__ void k_sub(double2* __restrict__ o, const double2* __restrict__ x,
                      const double2* __restrict__ y, long N) {
    GRID_STRIDE(n, N) o[n] = csub(x[n], y[n]);
}
__global__ void k_scale(double2* __restrict__ o, double a, const double2* __restrict__ x, long N) {
    GRID_STRIDE(n, N) o[n] = rmul(a, x[n]);
}
// x[n * stride] /= d for n < N (the divide loop of FFT1D::inverse)
__global__ void k_div_strided(double2* __restrict__ x, double d, long N, long stride) {
    GRID_STRIDE(n, N) {
        const double2 v = x[n * stride];
        x[n * stride] = mk<double>(v.x / d, v.y / d);
    }
}

// ---------------------------------------------------------------- reductions (deterministic two-pass)
static constexpr int kRedBlocks = 1024;

template <int MODE>
__global__ __launch_bounds__(kBlock) void k_reduce_partial(const double2* __restrict__ f, long N,
                                                           double* __restrict__ part) {
    double acc = 0;
    GRID_STRIDE(n, N) {
        const double2 v = f[n];
        if (MODE == 0)
            acc += hypot(v.x, v.y);
        else
            acc += v.x * v.x + v.y * v.y;
    }
    __shared__ double s[kBlock];
    s[threadIdx.x] = acc;
    __syncthreads();
    for (int w = kBlock / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = s[0];
}

__global__ __launch_bounds__(kBlock) void k_reduce_final(const double* __restrict__ part, int np,
                                                         double* __restrict__ out) {
    __shared__ double s[kBlock];
    double acc = 0;
    for (int i = threadIdx.x; i < np; i += kBlock) acc += part[i];
    s[threadIdx.x] = acc;
    __syncthreads();
    for (int w = kBlock / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = s[0];
}

int reduce_z(const void* f, long n, int mode, double* result, hipStream_t st) {
    double* buf = nullptr;
    int rc = pool_alloc((void**)&buf, sizeof(double) * (kRedBlocks + 1));
    if (rc) return rc;
    unsigned g = grid_for(n);
    if (g > (unsigned)kRedBlocks) g = kRedBlocks;
    if (mode == 0)
        hipLaunchKernelGGL(k_reduce_partial<0>, dim3(g), dim3(kBlock), 0, st, (const double2*)f, n, buf);
    else
        hipLaunchKernelGGL(k_reduce_partial<1>, dim3(g), dim3(kBlock), 0, st, (const double2*)f, n, buf);
    hipLaunchKernelGGL(k_reduce_final, dim3(1), dim3(kBlock), 0, st, buf, (int)g, buf + kRedBlocks);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(result, buf + kRedBlocks, sizeof(double), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    pool_free(buf);
    if (e != hipSuccess) return hip_fail(e, "reduce");
    return SH_OK;
}

// ---------------------------------------------------------------- initial conditions
__device__ __forceinline__ uint64_t splitmix(uint64_t seed, uint64_t c) {
    uint64_t z = seed + (c + 1) * 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}
__device__ __forceinline__ double u01(uint64_t seed, uint64_t c) {
    return (double)(splitmix(seed, c) >> 11) * 0x1.0p-53;
}

__global__ void k_prng_complex(double2* __restrict__ o, long N, uint64_t seed) {
    GRID_STRIDE(n, N) {
        o[n] = mk<double>(2.0 * u01(seed, 2 * (uint64_t)n) - 1.0, 2.0 * u01(seed, 2 * (uint64_t)n + 1) - 1.0);
    }
}

__global__ void k_taylor_green(double2* __restrict__ v, Dims d) {
    const long N = d.n();
    GRID_STRIDE(n, N) {
        const int i = (int)(n % d.nx);
        const long r = n / d.nx;
        const int j = (int)(r % d.ny);
        const int k = (int)(r / d.ny);
        const double x = (2. * M_PI / d.nx) * i, y = (2. * M_PI / d.ny) * j, z = (2. * M_PI / d.nz) * k;
        v[n] = mk<double>(sin(x) * cos(y) * cos(z), 0.0);
        v[N + n] = mk<double>(-cos(x) * sin(y) * cos(z), 0.0);
        v[2 * N + n] = mk<double>(0.0, 0.0);
    }
}

// unsymmetrised projected random mode m, component values u[3]
__device__ void solenoidal_mode(long m, const Dims& d, uint64_t seed, double k0, double2 u[3]) {
    double kx, ky, kz;
    wavenumber(m, d, kx, ky, kz);
    const double q = kx * kx + ky * ky + kz * kz;
    if (q <= 0) {
        u[0] = u[1] = u[2] = mk<double>(0, 0);
        return;
    }
    const double km = sqrt(q);
    const double E = pow(km, 4.0) * exp(-2.0 * (km / k0) * (km / k0));
    const double amp = sqrt(E / (4.0 * M_PI * q));
    double ur[3], ui[3];
    for (int c = 0; c < 3; c++) {
        const double ph = 2.0 * M_PI * u01(seed, 3 * (uint64_t)m + c);
        ur[c] = amp * cos(ph);
        ui[c] = amp * sin(ph);
    }
    const double kr = kx * ur[0] + ky * ur[1] + kz * ur[2];
    const double ki = kx * ui[0] + ky * ui[1] + kz * ui[2];
    const double inv = 1.0 / q;
    const double kk[3] = {kx, ky, kz};
    for (int c = 0; c < 3; c++) u[c] = mk<double>(ur[c] - kk[c] * (kr * inv), ui[c] - kk[c] * (ki * inv));
}

__global__ void k_solenoidal(double2* __restrict__ v, Dims d, uint64_t seed, double k0) {
    const long N = d.n();
    GRID_STRIDE(n, N) {
        const int i = (int)(n % d.nx);
        const long r = n / d.nx;
        const int j = (int)(r % d.ny);
        const int k = (int)(r / d.ny);
        const bool zero = (i == d.nx / 2) || (j == d.ny / 2) || (k == d.nz / 2) || n == 0;
        double2 a[3], b[3];
        solenoidal_mode(n, d, seed, k0, a);
        const long mn = (long)((d.nx - i) % d.nx) + (long)d.nx * ((d.ny - j) % d.ny) +
                        (long)d.nx * d.ny * ((d.nz - k) % d.nz);
        solenoidal_mode(mn, d, seed, k0, b);
        for (int c = 0; c < 3; c++)
            v[c * N + n] = zero ? mk<double>(0, 0) : mk<double>(0.5 * (a[c].x + b[c].x), 0.5 * (a[c].y - b[c].y));
    }
}

// Shell spectrum, deterministic (the same bits on every call): each wave
// walks 64-element chunks of the field; within a chunk the lanes of each
// distinct shell are summed by an xor butterfly (identical on every lane, as
// a + b == b + a) and lane 0 adds that sum to the wave's own LDS histogram, so
// every bin accumulates in a fixed order.  Entries nbins, nbins + 1 of a row
// hold the totals of e and |k|^2 e.  The block adds its waves' rows in wave
// order into its partial row; k_spectrum_final adds the rows in block order.
static constexpr int kSpecWaves = kBlock / 64;

__device__ __forceinline__ double wave_sum(double x) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
    return x;
}

__global__ __launch_bounds__(kBlock) void k_spectrum_partial(const double2* __restrict__ v, Dims d,
                                                             double norm, int nbins,
                                                             double* __restrict__ part) {
    extern __shared__ double sh[];  // kSpecWaves rows of nbins + 2
    const int W = nbins + 2;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    double* hist = sh + wave * W;
    for (int b = threadIdx.x; b < kSpecWaves * W; b += blockDim.x) sh[b] = 0;
    __syncthreads();
    const long N = d.n();
    const long stride = (long)gridDim.x * blockDim.x;
    double e_acc = 0, d_acc = 0;
    for (long base = (long)blockIdx.x * blockDim.x + 64 * wave; base < N; base += stride) {
        const long n = base + lane;
        double e = 0;
        int bin = -1;
        if (n < N) {
            double k0, k1, k2;
            wavenumber(n, d, k0, k1, k2);
            const double q = k0 * k0 + k1 * k1 + k2 * k2;
            const double2 a = v[n], b = v[N + n], c = v[2 * N + n];
            e = 0.5 * (a.x * a.x + a.y * a.y + b.x * b.x + b.y * b.y + c.x * c.x + c.y * c.y) * norm;
            e_acc += e;
            d_acc += q * e;
            const int sb = (int)floor(sqrt(q) + 0.5);
            if (sb < nbins) bin = sb;
        }
        unsigned long long pend = __ballot(bin >= 0);
        while (pend) {
            const int b = __shfl(bin, __ffsll(pend) - 1, 64);
            const double x = wave_sum(bin == b ? e : 0.0);
            if (lane == 0) hist[b] += x;
            pend &= ~__ballot(bin == b);
        }
    }
    e_acc = wave_sum(e_acc);
    d_acc = wave_sum(d_acc);
    if (lane == 0) {
        hist[nbins] = e_acc;
        hist[nbins + 1] = d_acc;
    }
    __syncthreads();
    for (int b = threadIdx.x; b < W; b += blockDim.x) {
        double acc = 0;
        for (int w = 0; w < kSpecWaves; w++) acc += sh[w * W + b];
        part[(long)blockIdx.x * W + b] = acc;
    }
}

__global__ __launch_bounds__(kBlock) void k_spectrum_final(const double* __restrict__ part, int np, int W,
                                                           double* __restrict__ out) {
    for (int b = threadIdx.x; b < W; b += blockDim.x) {
        double acc = 0;
        for (int r = 0; r < np; r++) acc += part[(long)r * W + b];
        out[b] = acc;
    }
}

// per-block partial sums of |k x u|^2 and |k . u|^2 (curl and divergence of
// src/Vector.cpp:92-138 without the factor i, which has modulus 1)
__global__ __launch_bounds__(kBlock) void k_vort_div_partial(const double2* __restrict__ v, Dims d,
                                                             double* __restrict__ part) {
    __shared__ double sw[kBlock], sd[kBlock];
    const long N = d.n();
    double w2 = 0, d2 = 0;
    GRID_STRIDE(n, N) {
        double k[3];
        wavenumber(n, d, k[0], k[1], k[2]);
        const double2 u[3] = {v[n], v[N + n], v[2 * N + n]};
        for (int c = 0; c < 3; c++) {
            const int a = (c + 1) % 3, b = (c + 2) % 3;
            const double cr = k[a] * u[b].x - k[b] * u[a].x;
            const double ci = k[a] * u[b].y - k[b] * u[a].y;
            w2 += cr * cr + ci * ci;
        }
        const double dr = k[0] * u[0].x + k[1] * u[1].x + k[2] * u[2].x;
        const double di = k[0] * u[0].y + k[1] * u[1].y + k[2] * u[2].y;
        d2 += dr * dr + di * di;
    }
    sw[threadIdx.x] = w2;
    sd[threadIdx.x] = d2;
    __syncthreads();
    for (int s = kBlock / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            sw[threadIdx.x] += sw[threadIdx.x + s];
            sd[threadIdx.x] += sd[threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = sw[0];
        part[2 * blockIdx.x + 1] = sd[0];
    }
}

// Gaussian-vortex vorticity w_z on x_i = 2 pi i / N (exec/ns.cpp:84-112 set 0,
// exec/ns2.cpp:84-112 set 1); the x and y components are zero
__device__ __forceinline__ double gauss2(double x, double y, double xc, double yc, double width) {
    return exp(-((x - xc) * (x - xc)) / width) * exp(-((y - yc) * (y - yc)) / width);
}

__global__ void k_vortex_vorticity(double2* __restrict__ w, Dims d, int set) {
    const long N = d.n();
    const double width = M_PI / 8, mag = 50, dist = M_PI / 2;
    GRID_STRIDE(n, N) {
        const int i = (int)(n % d.nx);
        const int j = (int)((n / d.nx) % d.ny);
        const double x = (2. * M_PI / d.nx) * i, y = (2. * M_PI / d.ny) * j;
        double wz;
        if (set == 0)
            wz = mag * gauss2(x, y, M_PI - dist, M_PI - dist, width) +
                 mag * gauss2(x, y, M_PI + dist, M_PI - dist, width) +
                 mag * gauss2(x, y, M_PI, M_PI + dist, width);
        else
            wz = mag * gauss2(x, y, M_PI - dist, M_PI, width) + mag * gauss2(x, y, M_PI + dist, M_PI, width);
        w[n] = mk<double>(0.0, 0.0);
        w[N + n] = mk<double>(0.0, 0.0);
        w[2 * N + n] = mk<double>(wz, 0.0);
    }
}

// psi = -w_z/(kx^2+ky^2) where (kx, ky) != 0, else 0; V0 = (-i ky psi, i kx psi, 0)
__global__ void k_vortex_velocity(const double2* __restrict__ wz, double2* __restrict__ v, Dims d) {
    const long N = d.n();
    GRID_STRIDE(n, N) {
        double k0, k1, k2;
        wavenumber(n, d, k0, k1, k2);
        double2 psi = mk<double>(0.0, 0.0);
        if (k0 != 0 || k1 != 0) psi = rmul<double>(-1. / (k0 * k0 + k1 * k1), wz[n]);
        const double2 gy = imul<double>(k1, psi);
        v[n] = mk<double>(-1. * gy.x, -1. * gy.y);
        v[N + n] = imul<double>(k0, psi);
        v[2 * N + n] = mk<double>(0.0, 0.0);
    }
}

static int check_dims(const Dims& d, const char* who) {
    if (d.nx <= 0 || d.ny <= 0 || d.nz <= 0) {
        set_error(std::string(who) + ": invalid dimensions");
        return SH_EINVAL;
    }
    return ensure_init();
}

// spectrum launch shape: depends on the field size only, so the summation
// order (and the result) is the same on every device and call
static constexpr int kSpecMaxBins = 2048;
static unsigned spectrum_grid(long n) {
    unsigned g = grid_for(n);
    return g > (unsigned)kRedBlocks ? (unsigned)kRedBlocks : g;
}

// partial rows into buf[0 .. g*W), their sum into buf[g*W ..) and then into
// host[0 .. W); synchronous
static int spectrum_run(const double2* v3, const Dims& d, int nbins, double* buf, unsigned g,
                        double* host, hipStream_t st) {
    const int W = nbins + 2;
    const double norm = 1.0 / ((double)d.n() * (double)d.n());
    hipLaunchKernelGGL(k_spectrum_partial, dim3(g), dim3(kBlock), sizeof(double) * kSpecWaves * W, st, v3, d,
                       norm, nbins, buf);
    hipLaunchKernelGGL(k_spectrum_final, dim3(1), dim3(kBlock), 0, st, (const double*)buf, (int)g, W,
                       buf + (size_t)g * W);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess)
        e = hipMemcpyAsync(host, buf + (size_t)g * W, sizeof(double) * W, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) return hip_fail(e, "sh_spectrum_z");
    return SH_OK;
}

static void spectrum_outputs(const double* h, int nbins, double nu, double* Ek, double* energy,
                             double* dissipation) {
    for (int b = 0; b < nbins; b++) Ek[b] = h[b];
    if (energy) *energy = h[nbins];
    if (dissipation) *dissipation = 2.0 * nu * h[nbins + 1];
}

// ---------------------------------------------------------------- spectrum without allocation
// The partial rows of sh_spectrum_z_noalloc live in `scratch`, a file-scope
// device array (each device holds its own copy until the process ends).  Host
// code never uses the name `scratch` as an address: its device address is
// taken once per device with hipGetSymbolAddress and kept in g_scratch_dev.
// Every entry of the rows a call reads is written by that call's kernels
// first, so nothing needs clearing between calls; calls that overlap on one
// device race on the array.
static constexpr int kSpecNoallocMaxBins = 510;
static constexpr size_t kSpecScratchDoubles = (size_t)(kRedBlocks + 1) * (kSpecNoallocMaxBins + 2);
static __device__ __align__(256) double scratch[kSpecScratchDoubles];
static std::mutex g_scratch_mu;
static std::vector<double*> g_scratch_dev;

// the current device's address of `scratch` (taken on first use per device)
static int scratch_address(double** out) {
    int dev = 0;
    SH_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if ((int)g_scratch_dev.size() <= dev) g_scratch_dev.resize(dev + 1, nullptr);
    if (!g_scratch_dev[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) return hip_fail(e, "spectrum scratch: hipGetSymbolAddress");
        g_scratch_dev[dev] = (double*)p;
    }
    *out = g_scratch_dev[dev];
    return SH_OK;
}

// sh_init's hook: take the address on the device being initialised
int spectrum_scratch_init() {
    double* p = nullptr;
    return scratch_address(&p);
}

}  // namespace sh

using namespace sh;

#define SH_REQUIRE(cond, who)                              \
    do {                                                   \
        if (!(cond)) {                                     \
            set_error(std::string(who) + ": null argument"); \
            return SH_EINVAL;                              \
        }                                                  \
    } while (0)

extern "C" {

int sh_zero_high_z(void* f, int nx, int ny, int nz, void* stream) {
    Dims d{nx, ny, nz};
    SH_REQUIRE(f, "sh_zero_high_z");
    int rc = check_dims(d, "sh_zero_high_z");
    if (rc) return rc;
    hipLaunchKernelGGL(k_zero_high, dim3(grid_for(d.n())), dim3(kBlock), 0, as_stream(stream), (double2*)f, d);
    SH_LAUNCH_CHECK("k_zero_high");
    return SH_OK;
}

int sh_gradient_z(const void* f, void* vx, void* vy, void* vz, int nx, int ny, int nz, void* stream) {
    Dims d{nx, ny, nz};
    SH_REQUIRE(f && vx && vy && vz, "sh_gradient_z");
    int rc = check_dims(d, "sh_gradient_z");
    if (rc) return rc;
    hipLaunchKernelGGL(k_gradient, dim3(grid_for(d.n())), dim3(kBlock), 0, as_stream(stream),
                       (const double2*)f, (double2*)vx, (double2*)vy, (double2*)vz, d);
    SH_LAUNCH_CHECK("k_gradient");
    return SH_OK;
}

int sh_laplacian_z(const void* f, void* out, int nx, int ny, int nz, void* stream) {
    Dims d{nx, ny, nz};
    SH_REQUIRE(f && out, "sh_laplacian_z");
    int rc = check_dims(d, "sh_laplacian_z");
    if (rc) return rc;
    hipLaunchKernelGGL(k_laplacian, dim3(grid_for(d.n())), dim3(kBlock), 0, as_stream(stream),
                       (const double2*)f, (double2*)out, d);
    SH_LAUNCH_CHECK("k_laplacian");
    return SH_OK;
}

int sh_laplacian_inverse_z(void* f, double a, double b, int nx, int ny, int nz, void* stream) {
    Dims d{nx, ny, nz};
    SH_REQUIRE(f, "sh_laplacian_inverse_z");
    int rc = check_dims(d, "sh_laplacian_inverse_z");
    if (rc) return rc;
    hipLaunchKernelGGL(k_laplacian_inverse, dim3(grid_for(d.n())), dim3(kBlock), 0, as_stream(stream),
                       (double2*)f, a, b, d);
    SH_LAUNCH_CHECK("k_laplacian_inverse");
    return SH_OK;
}

int sh_divergence_z(const void* vx, const void* vy, const void* vz, void* out, int nx, int ny, int nz,
                    void* stream) {
    Dims d{nx, ny, nz};
    SH_REQUIRE(vx && vy && vz && out, "sh_divergence_z");
    int rc = check_dims(d, "sh_divergence_z");
    if (rc) return rc;
    hipLaunchKernelGGL(k_divergence, dim3(grid_for(d.n())), dim3(kBlock), 0, as_stream(stream),
                       (const double2*)vx, (const double2*)vy, (const double2*)vz, (double2*)out, d);
    SH_LAUNCH_CHECK("k_divergence");
    return SH_OK;
}

int sh_curl_z(const void* vx, const void* vy, const void* vz, void* ox, void* oy, void* oz, int nx,
              int ny, int nz, void* stream) {
    Dims d{nx, ny, nz};
    SH_REQUIRE(vx && vy && vz && ox && oy && oz, "sh_curl_z");
    int rc = check_dims(d, "sh_curl_z");
    if (rc) return rc;
    hipLaunchKernelGGL(k_curl, dim3(grid_for(d.n())), dim3(kBlock), 0, as_stream(stream),
                       (const double2*)vx, (const double2*)vy, (const double2*)vz, (double2*)ox,
                       (double2*)oy, (double2*)oz, d);
    SH_LAUNCH_CHECK("k_curl");
    return SH_OK;
}

int sh_vlaplacian_z(const void* vx, const void* vy, const void* vz, void* ox, void* oy, void* oz,
                    int nx, int ny, int nz, void* stream) {
    Dims d{nx, ny, nz};
    SH_REQUIRE(vx && vy && vz && ox && oy && oz, "sh_vlaplacian_z");
    int rc = check_dims(d, "sh_vlaplacian_z");
    if (rc) return rc;
    hipLaunchKernelGGL(k_vlaplacian, dim3(grid_for(d.n())), dim3(kBlock), 0, as_stream(stream),
                       (const double2*)vx, (const double2*)vy, (const double2*)vz, (double2*)ox,
                       (double2*)oy, (double2*)oz, d);
    SH_LAUNCH_CHECK("k_vlaplacian");
    return SH_OK;
}

int sh_cross_pointwise_z(const void* ux, const void* uy, const void* uz, const void* wx,
                         const void* wy, const void* wz, void* ox, void* oy, void* oz, long n,
                         void* stream) {
    SH_REQUIRE(ux && uy && uz && wx && wy && wz && ox && oy && oz, "sh_cross_pointwise_z");
    if (n < 0) return SH_EINVAL;
    if (n == 0) return SH_OK;
    int rc = ensure_init();
    if (rc) return rc;
    hipLaunchKernelGGL(k_cross_pw, dim3(grid_for(n)), dim3(kBlock), 0, as_stream(stream),
                       (const double2*)ux, (const double2*)uy, (const double2*)uz, (const double2*)wx,
                       (const double2*)wy, (const double2*)wz, (double2*)ox, (double2*)oy,
                       (double2*)oz, n);
    SH_LAUNCH_CHECK("k_cross_pw");
    return SH_OK;
}

int sh_cross_z(const void* a3, const void* b3, void* out3, void* work6, int nx, int ny, int nz,
               void* stream) {
    Dims d{nx, ny, nz};
    SH_REQUIRE(a3 && b3 && out3 && work6, "sh_cross_z");
    int rc = check_dims(d, "sh_cross_z");
    if (rc) return rc;
    hipStream_t st = as_stream(stream);
    const long N = d.n();
    double2* w = (double2*)work6;
    // deep copies of both operands (src/Vector.cpp:75-76), then 6 inverse transforms
    SH_HIP(hipMemcpyAsync(w, a3, sizeof(double2) * 3 * N, hipMemcpyDeviceToDevice, st));
    SH_HIP(hipMemcpyAsync(w + 3 * N, b3, sizeof(double2) * 3 * N, hipMemcpyDeviceToDevice, st));
    rc = fft3d<double>(w, 6, N, nx, ny, nz, +1, SH_NORMALIZE, stream);
    if (rc) return rc;
    double2* o = (double2*)out3;
    hipLaunchKernelGGL(k_cross_pw, dim3(grid_for(N)), dim3(kBlock), 0, st, w, w + N, w + 2 * N,
                       w + 3 * N, w + 4 * N, w + 5 * N, o, o + N, o + 2 * N, N);
    SH_LAUNCH_CHECK("k_cross_pw");
    return fft3d<double>(o, 3, N, nx, ny, nz, -1, SH_ZERO_NYQUIST, stream);
}

int sh_mul_z(void* out, const void* x, const void* y, long n, void* stream) {
    SH_REQUIRE(out && x && y, "sh_mul_z");
    if (n <= 0) return n == 0 ? SH_OK : SH_EINVAL;
    hipLaunchKernelGGL(k_mul, dim3(grid_for(n)), dim3(kBlock), 0, as_stream(stream), (double2*)out,
                       (const double2*)x, (const double2*)y, n);
    SH_LAUNCH_CHECK("k_mul");
    return SH_OK;
}

int sh_add_z(void* out, const void* x, const void* y, long n, void* stream) {
    SH_REQUIRE(out && x && y, "sh_add_z");
    if (n <= 0) return n == 0 ? SH_OK : SH_EINVAL;
    hipLaunchKernelGGL(k_add, dim3(grid_for(n)), dim3(kBlock), 0, as_stream(stream), (double2*)out,
                       (const double2*)x, (const double2*)y, n);
    SH_LAUNCH_CHECK("k_add");
    return SH_OK;
}

int sh_sub_z(void* out, const void* x, const void* y, long n, void* stream) {
    SH_REQUIRE(out && x && y, "sh_sub_z");
    if (n <= 0) return n == 0 ? SH_OK : SH_EINVAL;
    hipLaunchKernelGGL(k_sub, dim3(grid_for(n)), dim3(kBlock), 0, as_stream(stream), (double2*)out,
                       (const double2*)x, (const double2*)y, n);
    SH_LAUNCH_CHECK("k_sub");
    return SH_OK;
}

int sh_scale_z(void* out, double alpha, const void* x, long n, void* stream) {
    SH_REQUIRE(out && x, "sh_scale_z");
    if (n <= 0) return n == 0 ? SH_OK : SH_EINVAL;
    hipLaunchKernelGGL(k_scale, dim3(grid_for(n)), dim3(kBlock), 0, as_stream(stream), (double2*)out,
                       alpha, (const double2*)x, n);
    SH_LAUNCH_CHECK("k_scale");
    return SH_OK;
}

int sh_div_strided_z(void* x, double divisor, long n, long stride, void* stream) {
    SH_REQUIRE(x, "sh_div_strided_z");
    if (n < 0 || stride < 1) {
        set_error("sh_div_strided_z: n >= 0 and stride >= 1");
        return SH_EINVAL;
    }
    if (n == 0) return SH_OK;
    int rc = ensure_init();
    if (rc) return rc;
    hipLaunchKernelGGL(k_div_strided, dim3(grid_for(n)), dim3(kBlock), 0, as_stream(stream), (double2*)x,
                       divisor, n, stride);
    SH_LAUNCH_CHECK("k_div_strided");
    return SH_OK;
}

int sh_average_z(const void* f, long n, double* result, void* stream) {
    SH_REQUIRE(f && result, "sh_average_z");
    if (n <= 0) {
        set_error("sh_average_z: empty field");
        return SH_EINVAL;
    }
    int rc = ensure_init();
    if (rc) return rc;
    double s = 0;
    rc = reduce_z(f, n, 0, &s, as_stream(stream));
    if (rc) return rc;
    *result = s / (double)n;
    return SH_OK;
}

int sh_sumsq_z(const void* f, long n, double* result, void* stream) {
    SH_REQUIRE(f && result, "sh_sumsq_z");
    if (n <= 0) {
        *result = 0;
        return n == 0 ? SH_OK : SH_EINVAL;
    }
    int rc = ensure_init();
    if (rc) return rc;
    return reduce_z(f, n, 1, result, as_stream(stream));
}

int sh_prng_complex_z(void* out, long n, uint64_t seed, void* stream) {
    SH_REQUIRE(out, "sh_prng_complex_z");
    if (n <= 0) return n == 0 ? SH_OK : SH_EINVAL;
    int rc = ensure_init();
    if (rc) return rc;
    hipLaunchKernelGGL(k_prng_complex, dim3(grid_for(n)), dim3(kBlock), 0, as_stream(stream),
                       (double2*)out, n, seed);
    SH_LAUNCH_CHECK("k_prng_complex");
    return SH_OK;
}

int sh_taylor_green_physical_z(void* v3, int nx, int ny, int nz, void* stream) {
    Dims d{nx, ny, nz};
    SH_REQUIRE(v3, "sh_taylor_green_physical_z");
    int rc = check_dims(d, "sh_taylor_green_physical_z");
    if (rc) return rc;
    hipLaunchKernelGGL(k_taylor_green, dim3(grid_for(d.n())), dim3(kBlock), 0, as_stream(stream),
                       (double2*)v3, d);
    SH_LAUNCH_CHECK("k_taylor_green");
    return SH_OK;
}

int sh_solenoidal_ic_z(void* v3, int nx, int ny, int nz, uint64_t seed, double k0, double urms,
                       void* stream) {
    Dims d{nx, ny, nz};
    SH_REQUIRE(v3, "sh_solenoidal_ic_z");
    int rc = check_dims(d, "sh_solenoidal_ic_z");
    if (rc) return rc;
    hipStream_t st = as_stream(stream);
    const long N = d.n();
    hipLaunchKernelGGL(k_solenoidal, dim3(grid_for(N)), dim3(kBlock), 0, st, (double2*)v3, d, seed, k0);
    SH_LAUNCH_CHECK("k_solenoidal");
    double e2 = 0;
    rc = reduce_z(v3, 3 * N, 1, &e2, st);
    if (rc) return rc;
    e2 /= (double)N * (double)N;
    if (!(e2 > 0)) {
        set_error("sh_solenoidal_ic_z: zero-energy field (grid too small for the spectrum?)");
        return SH_EINVAL;
    }
    const double s = urms / sqrt(e2 / 3.0);
    return sh_scale_z(v3, s, v3, 3 * N, stream);
}

int sh_spectrum_z(const void* v3, int nx, int ny, int nz, double nu, double* Ek, int nbins,
                  double* energy, double* dissipation, void* stream) {
    Dims d{nx, ny, nz};
    SH_REQUIRE(v3, "sh_spectrum_z");
    if (nbins < 0 || nbins > kSpecMaxBins || (nbins > 0 && !Ek)) {
        set_error("sh_spectrum_z: nbins must be 0.." + std::to_string(kSpecMaxBins) + " (Ek non-null if > 0)");
        return SH_EINVAL;
    }
    int rc = check_dims(d, "sh_spectrum_z");
    if (rc) return rc;
    hipStream_t st = as_stream(stream);
    const unsigned g = spectrum_grid(d.n());
    const int W = nbins + 2;
    double* buf = nullptr;
    rc = pool_alloc((void**)&buf, sizeof(double) * ((size_t)g + 1) * W);
    if (rc) return rc;
    std::vector<double> h(W);
    rc = spectrum_run((const double2*)v3, d, nbins, buf, g, h.data(), st);
    pool_free(buf);
    if (rc) return rc;
    spectrum_outputs(h.data(), nbins, nu, Ek, energy, dissipation);
    return SH_OK;
}

int sh_spectrum_z_noalloc(const void* v3, int nx, int ny, int nz, double nu, double* Ek, int nbins,
                          double* energy, double* dissipation, void* stream) {
    Dims d{nx, ny, nz};
    SH_REQUIRE(v3, "sh_spectrum_z_noalloc");
    if (nbins < 0 || nbins > kSpecNoallocMaxBins || (nbins > 0 && !Ek)) {
        set_error("sh_spectrum_z_noalloc: nbins must be 0.." + std::to_string(kSpecNoallocMaxBins) +
                  " (the size of its static scratch; Ek non-null if > 0)");
        return SH_EINVAL;
    }
    int rc = check_dims(d, "sh_spectrum_z_noalloc");
    if (rc) return rc;
    double* buf = nullptr;
    rc = scratch_address(&buf);
    if (rc) return rc;
    const unsigned g = spectrum_grid(d.n());  // at most kRedBlocks partial rows
    double h[kSpecNoallocMaxBins + 2];
    rc = spectrum_run((const double2*)v3, d, nbins, buf, g, h, as_stream(stream));
    if (rc) return rc;
    spectrum_outputs(h, nbins, nu, Ek, energy, dissipation);
    return SH_OK;
}

int sh_vorticity_stats_z(const void* v3, int nx, int ny, int nz, double* enstrophy, double* divnorm,
                         void* stream) {
    Dims d{nx, ny, nz};
    SH_REQUIRE(v3, "sh_vorticity_stats_z");
    int rc = check_dims(d, "sh_vorticity_stats_z");
    if (rc) return rc;
    hipStream_t st = as_stream(stream);
    unsigned g = grid_for(d.n());
    if (g > (unsigned)kRedBlocks) g = kRedBlocks;
    double* part = nullptr;
    rc = pool_alloc((void**)&part, sizeof(double) * 2 * g);
    if (rc) return rc;
    hipLaunchKernelGGL(k_vort_div_partial, dim3(g), dim3(kBlock), 0, st, (const double2*)v3, d, part);
    std::vector<double> h(2 * g);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(h.data(), part, sizeof(double) * 2 * g, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    pool_free(part);
    if (e != hipSuccess) return hip_fail(e, "sh_vorticity_stats_z");
    // fixed-order host sum of the block partials: the result does not depend on scheduling
    double w2 = 0, d2 = 0;
    for (unsigned b = 0; b < g; b++) {
        w2 += h[2 * b];
        d2 += h[2 * b + 1];
    }
    const double nn = (double)d.n() * (double)d.n();
    if (enstrophy) *enstrophy = 0.5 * w2 / nn;
    if (divnorm) *divnorm = sqrt(d2 / nn);
    return SH_OK;
}

int sh_vortex_ic_z(void* V0, void* omega0, int nx, int ny, int nz, int set, void* stream) {
    Dims d{nx, ny, nz};
    SH_REQUIRE(V0 && omega0, "sh_vortex_ic_z");
    if (set != 0 && set != 1) {
        set_error("sh_vortex_ic_z: set must be 0 (exec/ns.cpp) or 1 (exec/ns2.cpp)");
        return SH_EINVAL;
    }
    int rc = check_dims(d, "sh_vortex_ic_z");
    if (rc) return rc;
    hipStream_t st = as_stream(stream);
    const long N = d.n();
    double2* w = (double2*)omega0;
    hipLaunchKernelGGL(k_vortex_vorticity, dim3(grid_for(N)), dim3(kBlock), 0, st, w, d, set);
    SH_LAUNCH_CHECK("k_vortex_vorticity");
    // omega0.forward(): the x and y components are zero and stay zero
    rc = fft3d<double>(w + 2 * N, 1, N, nx, ny, nz, -1, SH_ZERO_NYQUIST, stream);
    if (rc) return rc;
    hipLaunchKernelGGL(k_vortex_velocity, dim3(grid_for(N)), dim3(kBlock), 0, st, (const double2*)(w + 2 * N),
                       (double2*)V0, d);
    SH_LAUNCH_CHECK("k_vortex_velocity");
    return SH_OK;
}

}  // extern "C"
