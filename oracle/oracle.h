/*
 * oracle.h -- CPU restatement of the spectral_upcxx hot path (TEST INFRASTRUCTURE).
 *
 * This library is the checker for the HIP implementation and the timed CPU
 * baseline.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg load it; the product path never does.
 *
 * Data layout follows the reference exactly: complex<double> interleaved
 * (re, im), element (i, j, k) at n = i + Nx*j + Nx*Ny*k
 * (src/Scalar.cpp:682-690).  Single rank (p = 1): the slab "transposes" of
 * src/Scalar.cpp:446-626 reduce to the row/slab copies the reference does for
 * its own rank (memcpy at :496 and :590), and they are kept.
 *
 * Sign convention (src/FFT1D.cpp:72-101, src/Scalar.cpp:208-229):
 *   dir = -1 : X[q] = sum_r x[r] exp(-2 pi i r q / N)     (forward)
 *   dir = +1 : X[q] = sum_r x[r] exp(+2 pi i r q / N)     (unnormalised)
 */
#ifndef SPECTRAL_ORACLE_H
#define SPECTRAL_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- PRNG: counter-form splitmix64, U[0,1) = (z >> 11) * 2^-53 ---- */
uint64_t or_splitmix64(uint64_t seed, uint64_t counter);
void or_prng_uniform(uint64_t seed, uint64_t first, size_t n, double* out);
/* complex field, Re/Im i.i.d. U[-1,1]; element e uses draws 2e, 2e+1 */
void or_prng_complex(uint64_t seed, int nx, int ny, int nz, double* out);

/* ---- 1D / 3D FFT ---- */
/* in-place strided 1D FFT of length n (FFT1D::transform, src/FFT1D.cpp:95-101) */
int or_fft1d(double* x, int n, long stride, int dir);
/* virtual slab ranks of the 3D transform (src/Scalar.cpp:446-626): 1 (default)
 * runs every phase with all OpenMP threads; R > 1 runs R rank threads, each on
 * its own slab with the slab puts as copies (used when R divides Ny and Nz) */
void or_set_ranks(int ranks);
int or_get_ranks(void);
/* Scalar::transform(DIR) (src/Scalar.cpp:400-646); tmp has the size of f */
int or_transform3d(double* f, double* tmp, int nx, int ny, int nz, int dir);
/* Scalar::forward: transform(-1) + zeroHighModes (src/Scalar.cpp:208-215) */
int or_forward(double* f, double* tmp, int nx, int ny, int nz);
/* Scalar::inverse: transform(+1) then /= Nx*Ny*Nz (src/Scalar.cpp:217-229) */
int or_inverse(double* f, double* tmp, int nx, int ny, int nz);
/* Scalar::zeroHighModes (src/Scalar.cpp:336-370) */
void or_zero_high(double* f, int nx, int ny, int nz);
/* Scalar::freq for one axis (src/Scalar.cpp:198-206) */
double or_freq(int ind, int n);

/* ---- spectral operators (Fourier-space inputs) ---- */
void or_gradient(const double* s, double* vx, double* vy, double* vz, int nx, int ny, int nz);
void or_laplacian(const double* s, double* out, int nx, int ny, int nz);
void or_laplacian_inverse(double* s, double a, double b, int nx, int ny, int nz);
void or_divergence(const double* vx, const double* vy, const double* vz, double* out,
                   int nx, int ny, int nz);
void or_curl(const double* vx, const double* vy, const double* vz,
             double* ox, double* oy, double* oz, int nx, int ny, int nz);
void or_vlaplacian(const double* vx, const double* vy, const double* vz,
                   double* ox, double* oy, double* oz, int nx, int ny, int nz);
/* Vector::cross (src/Vector.cpp:66-90); a, b, out are 3 contiguous fields each.
 * work must hold 7 fields (6 physical copies + FFT tmp). */
int or_cross(const double* a, const double* b, double* out, double* work, int nx, int ny, int nz);
/* Scalar::average (src/Scalar.cpp:15-22) */
double or_average(const double* f, size_t n);

/* ---- Navier-Stokes driver (exec/ns.cpp:54-221) ---- */
typedef struct or_ns or_ns;
or_ns* or_ns_create(int nx, int ny, int nz, double nu, double K);
void or_ns_destroy(or_ns* s);
/* V0 (Fourier, 3 fields); omega0 = curl(V0) is formed here */
int or_ns_set_initial(or_ns* s, const double* V0);
/* Pi0, Forward-Euler V1, omega1, cross0 (exec/ns.cpp:132-153) */
int or_ns_startup(or_ns* s);
/* nsteps iterations of exec/ns.cpp:167-217 */
int or_ns_steps(or_ns* s, int nsteps);
/* timing helper: V1 = V0, omega1 = omega0, cross0 = 0 (no start-up) */
int or_ns_prime_for_timing(or_ns* s);
/* current velocity V1 (Fourier, 3 fields) */
void or_ns_get_velocity(const or_ns* s, double* out);
void or_ns_get_vorticity(const or_ns* s, double* out);
/* div.average() of the last step (exec/ns.cpp:209-210) */
double or_ns_div_average(const or_ns* s);
/* disable the cross product (Stokes limit) for closed-form checks */
void or_ns_set_nonlinear(or_ns* s, int on);
/* 0: exec/ns.cpp's Crank-Nicolson viscous term; 1: exec/ns2.cpp's backward Euler */
void or_ns_set_scheme(or_ns* s, int scheme);
/* replace omega0 = curl(V0) with the caller's initial vorticity (before startup) */
void or_ns_set_vorticity(or_ns* s, const double* omega0);

/* ---- initial conditions and diagnostics ---- */
int or_ic_taylor_green(double* V, double* tmp, int nx, int ny, int nz);
void or_ic_solenoidal(double* V, int nx, int ny, int nz, uint64_t seed, double k0, double urms);
/* kinetic energy 0.5 <|u|^2> from Fourier coefficients (numpy ifftn scaling) */
double or_energy(const double* V, int nx, int ny, int nz);
/* shell spectrum E(k), k = round(|k|), nbins entries */
void or_spectrum(const double* V, int nx, int ny, int nz, double* Ek, int nbins);
/* dissipation rate 2 nu sum |k|^2 e(k) (e as or_energy's per-mode terms) */
double or_dissipation(const double* V, int nx, int ny, int nz, double nu);
/* enstrophy 0.5 <|curl u|^2> and divergence norm sqrt(<|div u|^2>) (either pointer may be NULL) */
void or_vorticity_stats(const double* V, int nx, int ny, int nz, double* enstrophy, double* divnorm);
/* Gaussian vortices: set 0 exec/ns.cpp:84-131 (three), set 1 exec/ns2.cpp:84-130
 * (two); V0 and omega0 are 3 Fourier fields each, tmp one field */
int or_ic_vortices(double* V0, double* omega0, double* tmp, int nx, int ny, int nz, int set);

/* ---- marker particles (src/Particle.cpp) ---- */
void or_particles_seed(double* loc, long np, uint64_t seed);
void or_particles_move(double* loc, double* vel, double* energy, long np, const double* U,
                       int nx, int ny, int nz, double K);

int or_num_threads(void);

#ifdef __cplusplus
}
#endif
#endif
