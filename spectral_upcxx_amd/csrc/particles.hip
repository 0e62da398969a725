// particles.hip -- marker particles carried by the physical velocity field
// (src/Particle.cpp:5-55, driven once per step by exec/ns2.cpp:191-204).
//
// One thread per particle.  Each particle reads its 3 coordinates, gathers
// the real parts of the 8 cell corners of the 3 velocity fields (24 scattered
// 8-byte reads), and writes position, velocity and energy: a latency-bound
// gather, ~270 algorithmic bytes per particle, negligible next to the three
// inverse transforms that produce the field.
//
// Floating-point contraction is off in the move kernel and the operations are
// the reference's, in its order, so results are bit-identical to a scalar
// IEEE double run of src/Particle.cpp:12-46 (the trilinear sum is
// left-to-right, each weight product left-to-right).
#include <cmath>

#include "sh_common.hpp"

namespace sh {
namespace {

constexpr int kPBlock = 256;

__device__ __forceinline__ uint64_t pmix(uint64_t seed, uint64_t c) {
    uint64_t z = seed + (c + 1) * 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

// coordinate d of particle first + p = 2 pi U(seed, 3 (first + p) + d), U the 53-bit uniform
__global__ void k_particles_seed(double* __restrict__ loc, long n3, uint64_t seed, long first) {
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n3; i += (long)gridDim.x * blockDim.x)
        loc[i] = (double)(pmix(seed, (uint64_t)(3 * first + i)) >> 11) * 0x1.0p-53 * 2 * M_PI;
}

__global__ __launch_bounds__(kPBlock) void k_particles_move(double* __restrict__ loc, double* __restrict__ vel,
                                                            double* __restrict__ energy, long np,
                                                            const double2* __restrict__ ux,
                                                            const double2* __restrict__ uy,
                                                            const double2* __restrict__ uz, int nx, int ny,
                                                            int nz, double K) {
#pragma clang fp contract(off)
    const long dims[3] = {nx, ny, nz};
    for (long p = blockIdx.x * (long)blockDim.x + threadIdx.x; p < np; p += (long)gridDim.x * blockDim.x) {
        double x[3], rd[3];
        long ind[3];
        for (int d = 0; d < 3; d++) {
            x[d] = loc[3 * p + d];
            const double h = 2 * M_PI / dims[d];
            const double q = x[d] / h;
            // truncation as the reference's double -> INT conversion; the lower
            // clamp only guards positions outside [0, 2 pi)
            long i = q >= 0 ? (q < (double)dims[d] ? (long)q : dims[d]) : 0;
            if (i + 1 >= dims[d]) i = dims[d] - 2;
            ind[d] = i;
            rd[d] = (x[d] - (double)i * h) / h;
        }
        const long dy = nx, dz = (long)nx * ny;
        const long base = ind[0] + dy * ind[1] + dz * ind[2];
        double v[3];
        for (int d = 0; d < 3; d++) {
            const double2* f = (d == 0 ? ux : d == 1 ? uy : uz) + base;
            v[d] = f[0].x * (1 - rd[0]) * (1 - rd[1]) * (1 - rd[2]) +
                   f[1].x * rd[0] * (1 - rd[1]) * (1 - rd[2]) +
                   f[1 + dy].x * rd[0] * rd[1] * (1 - rd[2]) +
                   f[dy].x * (1 - rd[0]) * rd[1] * (1 - rd[2]) +
                   f[dz].x * (1 - rd[0]) * (1 - rd[1]) * (rd[2]) +
                   f[1 + dz].x * rd[0] * (1 - rd[1]) * (rd[2]) +
                   f[1 + dy + dz].x * rd[0] * rd[1] * (rd[2]) +
                   f[dy + dz].x * (1 - rd[0]) * rd[1] * (rd[2]);
        }
        for (int d = 0; d < 3; d++) {
            double y = x[d] + v[d] * K;
            // periodic wrap of src/Particle.cpp:38-45
            if (y >= 2 * M_PI)
                y = y - 2 * M_PI;
            else if (y <= 0)
                y = 2 * M_PI + y;
            loc[3 * p + d] = y;
            vel[3 * p + d] = v[d];
        }
        if (energy) {
            double e = 0;
            for (int d = 0; d < 3; d++) e += v[d] * v[d];
            energy[p] = .5 * e;
        }
    }
}

unsigned pgrid(long n) {
    long g = (n + kPBlock - 1) / kPBlock;
    if (g > 65536L * 4) g = 65536L * 4;
    return (unsigned)(g < 1 ? 1 : g);
}

}  // namespace
}  // namespace sh

using namespace sh;

extern "C" {

int sh_particles_seed(void* loc, long np, uint64_t seed, long first, void* stream) {
    if (np < 0 || first < 0 || (np > 0 && !loc)) {
        set_error("sh_particles_seed: invalid arguments");
        return SH_EINVAL;
    }
    if (np == 0) return SH_OK;
    int rc = ensure_init();
    if (rc) return rc;
    hipLaunchKernelGGL(k_particles_seed, dim3(pgrid(3 * np)), dim3(kPBlock), 0, as_stream(stream), (double*)loc,
                       3 * np, seed, first);
    SH_LAUNCH_CHECK("k_particles_seed");
    return SH_OK;
}

int sh_particles_move(void* loc, void* vel, void* energy, long np, const void* ux, const void* uy,
                      const void* uz, int nx, int ny, int nz, double K, void* stream) {
    if (np < 0 || (np > 0 && (!loc || !vel || !ux || !uy || !uz))) {
        set_error("sh_particles_move: invalid arguments");
        return SH_EINVAL;
    }
    if (nx < 2 || ny < 2 || nz < 2) {
        set_error("sh_particles_move: trilinear interpolation needs at least 2 points per axis");
        return SH_EINVAL;
    }
    if (np == 0) return SH_OK;
    int rc = ensure_init();
    if (rc) return rc;
    hipLaunchKernelGGL(k_particles_move, dim3(pgrid(np)), dim3(kPBlock), 0, as_stream(stream), (double*)loc,
                       (double*)vel, (double*)energy, np, (const double2*)ux, (const double2*)uy,
                       (const double2*)uz, nx, ny, nz, K);
    SH_LAUNCH_CHECK("k_particles_move");
    return SH_OK;
}

}  // extern "C"
