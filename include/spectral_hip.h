/*
 * spectral_hip.h -- C ABI of the MI355X (gfx950) spectral Navier-Stokes hot path.
 *
 * Plain C: device pointers, sizes and an opaque stream.  `stream` is a
 * hipStream_t passed as void* (NULL = the default stream).  Every entry point
 * returns 0 on success or a negative SH_E* code; sh_last_error() gives the
 * message of the last failure on the calling thread.  Launches are
 * stream-ordered and asynchronous unless the comment says "synchronous".
 *
 * Data layout everywhere is the reference's: interleaved complex
 * (complex<double> = "z", complex<float> = "c"), element (i, j, k) of an
 * Nx x Ny x Nz field at n = i + Nx*j + Nx*Ny*k (src/Scalar.cpp:682-690).
 * A "vector" argument is three such fields; a `v3` pointer means the three
 * fields are contiguous (component d at v3 + d*Nx*Ny*Nz elements).
 *
 * FFT lengths per axis: any N from 1 to 2048 whose prime factors are 2, 3,
 * 5, 7, 11, 13, and any other N up to 1024 (as FFTW plans any length,
 * src/FFT1D.cpp:14-54).  Compiled register/LDS plans: 8 10 12 14 16 20 24 28
 * 32 40 48 56 64 80 96 100 128 256 512 1024; the other smooth lengths run a
 * mixed-radix kernel with the radix list as an argument, the rest Bluestein's
 * chirp-z transform.  The fused solver schedule has kernels for 8 12 16 24 32
 * 48 64 96 128 256 512 per axis; the op-by-op schedule runs every length.
 * Sign convention (src/FFT1D.cpp:72-101): dir = -1 gives exp(-2 pi i jk/N)
 * (forward), dir = +1 gives exp(+2 pi i jk/N), unnormalised.
 */
#ifndef SPECTRAL_HIP_H
#define SPECTRAL_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    SH_OK = 0,
    SH_EINVAL = -1,      /* bad argument (null pointer, size, direction) */
    SH_EUNSUPPORTED = -2, /* FFT length (or fused-schedule grid) not supported */
    SH_EHIP = -3,        /* HIP runtime error */
    SH_ENOMEM = -4       /* device allocation failed */
};

/* precision tags */
enum { SH_F64 = 64, SH_F32 = 32 };

/* 3D transform flags */
enum {
    SH_ZERO_NYQUIST = 1, /* Scalar::zeroHighModes after the transform (src/Scalar.cpp:336-370) */
    SH_NORMALIZE = 2     /* divide by Nx*Ny*Nz (src/Scalar.cpp:223-227) */
};

/* ------------------------------------------------------------------ runtime */
/* Select the device and build per-device tables.  Optional: every entry
 * point initialises the current device on first use. */
int sh_init(int device);
const char* sh_last_error(void);
int sh_device_count(int* count);
/* Pooled device allocation (replaces upcxx::new_array/delete_array,
 * src/Scalar.cpp:95, 117, 38-40).  Freed blocks are cached for reuse. */
int sh_malloc(void** ptr, size_t bytes);
int sh_free(void* ptr);
/* Release every cached (free) pool block back to HIP. */
int sh_pool_trim(void);
/* synchronous copies */
int sh_memcpy_h2d(void* dst, const void* src, size_t bytes);
int sh_memcpy_d2h(void* dst, const void* src, size_t bytes);
/* stream-ordered device copy (the deep copies of src/Scalar.cpp:52-66) */
int sh_memcpy_d2d(void* dst, const void* src, size_t bytes, void* stream);
int sh_memset(void* dst, int value, size_t bytes, void* stream);
/* Streaming device copy (16-byte vectors per lane, nontemporal loads and
 * stores) used as the measured HBM bandwidth ceiling beside the kernels'
 * rooflines; kernel class "stream_copy" for sh_prof_query.  bytes a multiple
 * of 16, pointers 16-byte aligned; blocks = 0 launches one trip per lane,
 * otherwise that many 256-lane blocks loop over the buffer.  No counterpart
 * in the reference. */
int sh_stream_copy(void* dst, const void* src, size_t bytes, int blocks, void* stream);
int sh_stream_synchronize(void* stream);

/* ------------------------------------------------------------------ FFT */
/* Batched strided 1D C2C FFT in place: line b starts at data + b*dist and
 * holds n elements at `stride`.  Replaces FFT1D::transform(ptr, DIR)
 * (src/FFT1D.H:21-35, src/FFT1D.cpp:95-101), batched. */
int sh_fft1d_z2z(void* data, int n, long stride, long dist, long batch, int dir, void* stream);
int sh_fft1d_c2c(void* data, int n, long stride, long dist, long batch, int dir, void* stream);

/* 3D C2C FFT in place.  Replaces Scalar::transform(DIR) (src/Scalar.cpp:400-646):
 * the slab transposes of the reference are not needed with the whole field
 * in one device's memory.  flags: SH_ZERO_NYQUIST | SH_NORMALIZE.
 *   Scalar::forward() == sh_fft3d_z2z(f, .., -1, SH_ZERO_NYQUIST, ..)
 *   Scalar::inverse() == sh_fft3d_z2z(f, .., +1, SH_NORMALIZE, ..) */
int sh_fft3d_z2z(void* data, int nx, int ny, int nz, int dir, int flags, void* stream);
int sh_fft3d_c2c(void* data, int nx, int ny, int nz, int dir, int flags, void* stream);
/* the same transform on `count` fields spaced `field_stride` elements apart */
int sh_fft3d_many_z2z(void* data, int count, long field_stride, int nx, int ny, int nz, int dir,
                      int flags, void* stream);
int sh_fft3d_many_c2c(void* data, int count, long field_stride, int nx, int ny, int nz, int dir,
                      int flags, void* stream);

/* ------------------------------------------------------------------ spectral operators (complex128)
 * Fourier-space inputs; wavenumbers from Scalar::freq (src/Scalar.cpp:198-206). */
/* Scalar::zeroHighModes (src/Scalar.cpp:336-370) */
int sh_zero_high_z(void* f, int nx, int ny, int nz, void* stream);
/* Scalar::gradient: v_d = i k_d f (src/Scalar.cpp:245-276) */
int sh_gradient_z(const void* f, void* vx, void* vy, void* vz, int nx, int ny, int nz, void* stream);
/* Scalar::laplacian: -|k|^2 f (src/Scalar.cpp:278-306) */
int sh_laplacian_z(const void* f, void* out, int nx, int ny, int nz, void* stream);
/* Scalar::laplacian_inverse(a, b): f /= (a - b|k|^2) where |k|^2 != 0, in place
 * (src/Scalar.cpp:308-334) */
int sh_laplacian_inverse_z(void* f, double a, double b, int nx, int ny, int nz, void* stream);
/* Vector::divergence: sum_d i k_d v_d (src/Vector.cpp:124-153) */
int sh_divergence_z(const void* vx, const void* vy, const void* vz, void* out, int nx, int ny,
                    int nz, void* stream);
/* Vector::curl: i k x v (src/Vector.cpp:92-122) */
int sh_curl_z(const void* vx, const void* vy, const void* vz, void* ox, void* oy, void* oz,
              int nx, int ny, int nz, void* stream);
/* Vector::laplacian: -|k|^2 v_d (src/Vector.cpp:155-181) */
int sh_vlaplacian_z(const void* vx, const void* vy, const void* vz, void* ox, void* oy, void* oz,
                    int nx, int ny, int nz, void* stream);
/* Pointwise complex cross product of physical fields, out = u x w
 * (the loop of src/Vector.cpp:79-84). */
int sh_cross_pointwise_z(const void* ux, const void* uy, const void* uz, const void* wx,
                         const void* wy, const void* wz, void* ox, void* oy, void* oz, long n,
                         void* stream);
/* Vector::cross (src/Vector.cpp:66-90): out = forward(inverse(a) x inverse(b)).
 * a3, b3, out3: 3 contiguous Fourier fields; work: 6 fields of scratch. */
int sh_cross_z(const void* a3, const void* b3, void* out3, void* work6, int nx, int ny, int nz,
               void* stream);

/* ------------------------------------------------------------------ elementwise (Operators, src/Operators.cpp:9-99) */
/* out = x * y (Scalar * Scalar) */
int sh_mul_z(void* out, const void* x, const void* y, long n, void* stream);
/* out = x + y, out = x - y */
int sh_add_z(void* out, const void* x, const void* y, long n, void* stream);
int sh_sub_z(void* out, const void* x, const void* y, long n, void* stream);
/* out = alpha * x (double * Scalar / double * Vector) */
int sh_scale_z(void* out, double alpha, const void* x, long n, void* stream);
/* x[i*stride] /= divisor for i < n, in place: the divide loop of
 * FFT1D::inverse (src/FFT1D.cpp:59-66) as one strided launch */
int sh_div_strided_z(void* x, double divisor, long n, long stride, void* stream);
/* Scalar::average: mean |f| (src/Scalar.cpp:15-22).  Synchronous. */
int sh_average_z(const void* f, long n, double* result, void* stream);
/* sum of |f|^2 over n elements, FP64 accumulation.  Synchronous. */
int sh_sumsq_z(const void* f, long n, double* result, void* stream);

/* ------------------------------------------------------------------ initial conditions and diagnostics */
/* counter splitmix64 complex field, Re/Im i.i.d. U[-1,1] (element e: draws 2e, 2e+1) */
int sh_prng_complex_z(void* out, long n, uint64_t seed, void* stream);
/* Taylor-Green velocity (physical, complex128, 3 contiguous fields) on x_i = 2 pi i / N */
int sh_taylor_green_physical_z(void* v3, int nx, int ny, int nz, void* stream);
/* random solenoidal Fourier velocity, E(k) ~ k^4 exp(-2 (k/k0)^2), scaled to u_rms */
int sh_solenoidal_ic_z(void* v3, int nx, int ny, int nz, uint64_t seed, double k0, double urms,
                       void* stream);
/* shell energy spectrum of a Fourier velocity: Ek[b] = sum over the modes with
 * round(|k|) = b of e(k) = 0.5 |u(k)|^2 / (Nx Ny Nz)^2, b < nbins (at most
 * 2048; nbins may be 0).  Also the total energy sum e(k) = 0.5<|u|^2> and the
 * dissipation 2 nu sum |k|^2 e(k) if the pointers are non-null.  FP64 sums in
 * an order fixed by the grid size alone (no atomics): repeated calls give the
 * same bits.  Synchronous.  No counterpart in the reference (its diagnostic
 * is the divergence check of exec/ns.cpp:209-210). */
int sh_spectrum_z(const void* v3, int nx, int ny, int nz, double nu, double* Ek, int nbins,
                  double* energy, double* dissipation, void* stream);
/* sh_spectrum_z without any device allocation, for callers that cannot
 * allocate (code under graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  No counterpart in the reference.  The same kernels and
 * the same results, bit for bit, as sh_spectrum_z; the partial sums live in a
 * static device array of 1025 x 512 doubles (4 MiB per device, held until the
 * process ends), so nbins is at most 510 (SH_EINVAL beyond).  Calls that
 * overlap on one device race on that array.  Synchronous. */
int sh_spectrum_z_noalloc(const void* v3, int nx, int ny, int nz, double nu, double* Ek, int nbins,
                          double* energy, double* dissipation, void* stream);
/* enstrophy 0.5<|curl u|^2> and divergence norm sqrt(<|div u|^2>) of a Fourier
 * velocity, curl and divergence as src/Vector.cpp:92-138 (either output may
 * be NULL; FP64 block sums added in a fixed order).  Synchronous. */
int sh_vorticity_stats_z(const void* v3, int nx, int ny, int nz, double* enstrophy, double* divnorm,
                         void* stream);
/* Gaussian-vortex initial condition of the reference drivers: omega0 =
 * (0, 0, w_z), w_z = 50 sum_v exp(-(x-x_v)^2/w) exp(-(y-y_v)^2/w), w = pi/8,
 * on x_i = 2 pi i / N, forward()-ed (Nyquist zeroed); psi = -w_z/(kx^2+ky^2)
 * where (kx, ky) != 0; V0 = (-d_y psi, d_x psi, 0).  set 0: three vortices
 * (exec/ns.cpp:84-131); set 1: two vortices (exec/ns2.cpp:84-130).  V0 and
 * omega0 are 3 contiguous Fourier complex128 fields each. */
int sh_vortex_ic_z(void* V0, void* omega0, int nx, int ny, int nz, int set, void* stream);

/* ------------------------------------------------------------------ marker particles (src/Particle.cpp) */
/* Particle::Particle (src/Particle.cpp:5-10) for particles first ..
 * first+np-1: positions loc[np][3] in [0, 2 pi); the reference draws rand(),
 * here coordinate d of particle p is 2 pi U(seed, 3p + d) (splitmix64,
 * 53-bit).  Device memory. */
int sh_particles_seed(void* loc, long np, uint64_t seed, long first, void* stream);
/* Particle::move(velocity, K) (src/Particle.cpp:12-46) for np particles, as
 * exec/ns2.cpp:191-204 calls it once per step: trilinear interpolation of the
 * real parts of ux, uy, uz (the physical velocity components, complex128
 * fields as Vector::inverse leaves them) at each position, with the reference's cell
 * clamp (an index at dims-1 moves to dims-2), then loc += v K and the periodic
 * wrap.  vel[np][3] receives the interpolated velocity; energy[np] (may be
 * NULL) receives Particle::energy() = 0.5 |v|^2 (src/Particle.cpp:48-55).
 * Same IEEE operations in the same order, no contraction: bit-identical to a
 * scalar double run of the reference.  Needs nx, ny, nz >= 2. */
int sh_particles_move(void* loc, void* vel, void* energy, long np, const void* ux, const void* uy,
                      const void* uz, int nx, int ny, int nz, double K, void* stream);

/* ------------------------------------------------------------------ Navier-Stokes fractional step (exec/ns.cpp:132-217) */
typedef struct sh_ns_s* sh_ns_t;
/* precision: SH_F64 (plain method) or SH_F32 (FP32 solver; twiddles, CN
 * factors and reductions stay FP64).  fused: 1 = the fused kernel schedule,
 * 0 = one kernel per reference operation. */
int sh_ns_create(sh_ns_t* handle, int nx, int ny, int nz, double nu, double K, int precision,
                 int fused, void* stream);
int sh_ns_destroy(sh_ns_t handle);
/* V0: 3 contiguous Fourier fields, complex128 device memory */
int sh_ns_set_initial(sh_ns_t handle, const void* V0);
/* Replace omega0 = curl(V0), formed by sh_ns_set_initial, with the caller's
 * initial vorticity (3 Fourier fields) before sh_ns_startup: exec/ns.cpp:113-135
 * and exec/ns2.cpp:113-137 start from the vorticity V0 was built from, whose
 * mean mode curl(V0) lacks (see sh_vortex_ic_z). */
int sh_ns_set_vorticity(sh_ns_t handle, const void* omega0);
/* exec/ns.cpp:132-153: Pi0, Forward-Euler V1, omega1, cross0 */
int sh_ns_startup(sh_ns_t handle);
/* nsteps iterations of exec/ns.cpp:167-217 */
int sh_ns_steps(sh_ns_t handle, int nsteps);
/* current velocity V1 into 3 contiguous complex128 device fields.  These
 * three single-domain getters return SH_EINVAL on an ensemble handle (more
 * than one box): use the *_box calls there. */
int sh_ns_get_velocity(sh_ns_t handle, void* out_v3);
int sh_ns_get_vorticity(sh_ns_t handle, void* out_v3);
/* div.average() of the last step (exec/ns.cpp:209-210).  Synchronous. */
int sh_ns_div_average(sh_ns_t handle, double* result);
/* Viscous time discretisation: 0 (default) the Crank-Nicolson step of
 * exec/ns.cpp:174-187; 1 the backward-Euler step of exec/ns2.cpp:178-189
 * (lap1 = nu*K*V1.laplacian(), V = Vhalf divided by 1 + nu*K*|k|^2).  The
 * start-up is the same for both (exec/ns2.cpp:133-137).  Synchronous. */
int sh_ns_set_scheme(sh_ns_t handle, int scheme);
/* Stokes limit switch (nonlinear term off), for closed-form checks */
int sh_ns_set_nonlinear(sh_ns_t handle, int on);
/* Storage of the fused schedule.  1 (default): when the start-up state (V1,
 * cross0) is the transform of a real field -- Hermitian, f(-k) = conj f(k),
 * to 1e-12 relative, as every real initial velocity gives -- only the planes
 * kz < nz/2 are stepped (kz = nz/2 is zeroed by forward() anyway) and the z
 * transforms run as half-length complex ones on packed real data; otherwise
 * the full spectrum.  0: always the full complex spectrum, the reference's
 * C2C arithmetic (src/Scalar.cpp:400-646) operation for operation.  Both give
 * the same velocity to rounding on real fields; the half layout moves half
 * the bytes.  A body force is then read on the planes kz < nz/2 only (the
 * transform of a real force).  Takes effect at the next start-up. */
int sh_ns_set_layout(sh_ns_t handle, int mode);
/* 1 if the solver currently steps the half-spectrum layout, else 0 */
int sh_ns_get_layout(sh_ns_t handle, int* half);
/* Replaces nothing in the reference: the form of the fused state that the
 * last kernel E left.  0: V1 and cross0.  1: the merged AB2 history
 * (exec/ns.cpp:171-186 need V1 and cross0 only through Y = (V1 + lap1) -
 * K/2 cross0, so kernel E keeps Y and the velocity's x-inverse: 15 field
 * passes instead of 21).  2: the same with five emitted fields (the x-inverses
 * of V and of d/dx Vy, d/dx Vz; the next y pass forms the vorticity from them:
 * kernel E 14 passes, the y inverse 11 instead of 12).  The getters,
 * snapshots and schedule changes see (V1, cross0) in every form. */
int sh_ns_get_history_form(sh_ns_t handle, int* form);

/* Replaces nothing in the reference (graph capture is this implementation's):
 * puts the fused state in the form the current schedule steps -- the
 * x-inverse of V1 that the first kernel of a step reads, and the merged AB2
 * history when the chosen kernel E keeps it -- without advancing time, so
 * that a graph captured next records the merged kernels (a capture never
 * converts the state: a replay would convert it again).  The velocity is
 * unchanged to rounding.  Not inside a capture. */
int sh_ns_prepare(sh_ns_t handle);
/* Fused schedule only.  1: the y passes and the fused z kernel of a step run
 * as two independent halves of the x columns on the solver's stream and a
 * second (non-blocking, default-priority) stream of its own, joined before
 * the Fourier update, so that one half's transforms overlap the other's z
 * kernel; 0 (default): one launch each.  The same arithmetic either way.
 * sh_ns_autotune times both and keeps the faster. */
int sh_ns_set_overlap(sh_ns_t handle, int on);
/* device bytes held by the solver state */
int sh_ns_state_bytes(sh_ns_t handle, size_t* bytes);

/* ------------------------------------------------------------------ ensembles of equal boxes
 * No counterpart in the reference (one domain per run there).  nboxes boxes
 * of n^3 points share every launch of the fused step; box b has viscosity
 * nu[b] (host array), all share the time step K.  Boxes of different sizes
 * are separate handles created on one stream and captured into one graph. */
int sh_ns_create_ensemble(sh_ns_t* handle, int n, int nboxes, const double* nu, double K,
                          int precision, void* stream);
int sh_ns_box_count(sh_ns_t handle, int* count);
/* V0 of box `box` (3 contiguous complex128 Fourier fields), then that box's
 * start-up (exec/ns.cpp:132-153) */
int sh_ns_set_initial_box(sh_ns_t handle, int box, const void* V0);
int sh_ns_get_velocity_box(sh_ns_t handle, int box, void* out_v3);
int sh_ns_div_average_box(sh_ns_t handle, int box, double* result);
/* Box-filtered training snapshot of every box: physical velocity averaged
 * over ratio^3 cells.  out: float[nboxes][9][(n/ratio)^3] holding ubar_x,
 * ubar_y, ubar_z, tau_xx, tau_yy, tau_zz, tau_xy, tau_xz, tau_yz with
 * tau_ij = <u_i u_j> - ubar_i ubar_j.  The first call allocates scratch. */
int sh_ns_snapshot(sh_ns_t handle, int ratio, void* out);
/* Box-filtered nonlinear term of every box (target of the force-matching
 * closure loss): ratio^3 averages of the physical u x omega.
 * out: float[nboxes][3][(n/ratio)^3].  Shares the snapshot scratch. */
int sh_ns_snapshot_force(sh_ns_t handle, int ratio, void* out);
/* Body force added to the nonlinear term: each step uses cross1 + f in place
 * of cross1 (exec/ns.cpp:170-171), so f is Adams-Bashforth extrapolated with
 * it; the start-up adds f to both cross products it forms.  f: 3 Fourier
 * fields in the solver's precision (complex128, or complex64 for SH_F32),
 * laid out like the state (field d of box b at (d*nboxes + b)*N).  Its
 * contents are read at every step, so the caller may rewrite them between
 * steps.  NULL removes the force.  No counterpart in the reference. */
int sh_ns_set_forcing(sh_ns_t handle, const void* f3);
/* Dealiasing of the nonlinear term (with the body force): mode 1 drops every
 * mode with some |k_i| > N_i/3 (2/3 rule); mode 2 scales each mode by
 * prod_i exp(-36 (|k_i| / (N_i/2))^36) (exponential filter); mode 0, the
 * default, is the reference's scheme, which only zeroes the Nyquist modes.
 * No counterpart in the reference; used by the coarse-grid LES.  Synchronous. */
int sh_ns_set_dealias(sh_ns_t handle, int mode);

/* ------------------------------------------------------------------ streams and graphs
 * Default-priority non-blocking streams; stream capture into a hipGraph
 * (launch everything enqueued between begin and end as one graph). */
int sh_stream_create(void** stream);
int sh_stream_destroy(void* stream);
int sh_graph_begin(void* stream);
int sh_graph_end(void* stream, void** graph);
int sh_graph_launch(void* graph, void* stream);
int sh_graph_destroy(void* graph);

/* ------------------------------------------------------------------ schedule autotuner
 * Each pass role has a list of compiled radix/tile schedules that compute the
 * same transform; the tuner times them on the device and the dispatchers use
 * the fastest.  No counterpart in the reference (FFTW_MEASURE planning,
 * src/FFT1D.cpp:18-22, plays this part there).  Choices persist per process. */
enum {
    SH_ROLE_X = 0,          /* x lines of a 3D transform (contiguous pass) */
    SH_ROLE_Y = 1,          /* y lines (strided pass; also the solver's y passes) */
    SH_ROLE_Z = 2,          /* z lines (strided pass) */
    SH_ROLE_NS_CROSS = 3,   /* the solver's fused z / cross-product kernel */
    SH_ROLE_NS_FOURIER = 4, /* the solver's fused x / Fourier-update kernel */
    SH_ROLE_NS_CROSS_H = 5  /* the fused z kernel of the real-field layout (n = nz) */
};
/* time the x, y, z schedules for an nx x ny x nz transform.  Synchronous. */
int sh_autotune_fft3d(int nx, int ny, int nz, int precision, int reps, void* stream);
/* time the solver's schedules on its own grid (reps steps each) and restore
 * its state.  Synchronous. */
int sh_ns_autotune(sh_ns_t handle, int reps);
int sh_tune_variant_count(int precision, int n, int role, int* count);
/* force a schedule (0 = default) */
int sh_tune_set_variant(int precision, int n, int role, int variant);
int sh_tune_reset(void);
/* one line per tuned (precision, length, role): chosen schedule and its time */
int sh_tune_report(char* buf, size_t len);

/* ------------------------------------------------------------------ learned subgrid-stress closure
 * No counterpart in the reference.  An MLP 9 -> 64 -> 64 -> 6 (tanh hidden
 * layers) maps the filtered velocity-gradient tensor g_{3i+j} = d ubar_i/d x_j
 * to the stress components tau_xx, tau_yy, tau_zz, tau_xy, tau_xz, tau_yz.
 * Weights are FP32 master copies rounded to BF16 for the MFMA products,
 * which accumulate in FP32.  Inputs are multiplied by in_scale before the
 * network and outputs by out_scale after it. */
typedef struct sh_mlp_s* sh_mlp_t;
/* number of floats in the packed parameter vector: W1[16][64] (rows 9..15
 * zero), b1[64], W2[64][64], b2[64], W3[64][16] (columns 6..15 zero), b3[16] */
int sh_mlp_param_count(int* count);
/* Glorot-uniform weights from the splitmix64 stream of `seed`, zero biases,
 * unit scales */
int sh_mlp_create(sh_mlp_t* handle, uint64_t seed, void* stream);
int sh_mlp_destroy(sh_mlp_t handle);
int sh_mlp_set_params(sh_mlp_t handle, const float* host_params);
int sh_mlp_get_params(sh_mlp_t handle, float* host_params);
int sh_mlp_set_scaling(sh_mlp_t handle, const float* in_scale9, const float* out_scale6);
/* Y[f*ldy + s] = out_scale_f * net(in_scale * X[:, s])_f for s < n;
 * X: float device memory [9][ld], Y: float [6][ldy] */
int sh_mlp_forward(sh_mlp_t handle, const void* X, long ld, long n, void* Y, long ldy);
/* Gradient of the loss mean_{s<n, f<6} (net(X_s)_f - T[f*ld+s]/out_scale_f)^2
 * w.r.t. the packed parameters (padding entries zero) into grad_out (float
 * device memory, may be NULL); *loss (may be NULL; synchronous then) */
int sh_mlp_gradient(sh_mlp_t handle, const void* X, const void* T, long ld, long n, void* grad_out,
                    double* loss);
/* one Adam step (beta 0.9 / 0.999, eps 1e-8) on that gradient */
int sh_mlp_train_step(sh_mlp_t handle, const void* X, const void* T, long ld, long n, float lr,
                      double* loss);
/* Force-matching loss on one n^3 coarse field: with tau = net(X) (X: float
 * [9][n^3] features), f = -div tau as the LES applies it, and the target
 * force ft (3 Fourier complex64 fields), r = M P (f - ft) with P the
 * divergence-free projection and M zero on the Nyquist planes and at k = 0,
 * and times the coarse solver's dealiasing factor of mode `dealias`
 * (sh_ns_set_dealias); loss = weight * mean over modes of
 * w(k) |r(k)|^2 in Parseval's normalisation, w = 1 / (1 + |k|^2 / k_weight^2)
 * (k_weight = 0: w = 1, the mean of |r(x)|^2 over the grid); weight > 0,
 * e.g. the inverse of the target's own loss for a relative loss.  The
 * parameter gradient goes back through the divergence by its adjoint
 * (dL/dtau_ij = weight (2/n^3)(d_j s_i + d_i s_j) with s the inverse
 * transform of w r, halved on the diagonal) and through the network's
 * backward kernel. */
int sh_mlp_force_gradient(sh_mlp_t handle, int n, const void* X, const void* ft, int dealias,
                          double weight, double k_weight, void* grad_out, double* loss);
/* one Adam step on that gradient */
int sh_mlp_force_train_step(sh_mlp_t handle, int n, const void* X, const void* ft, int dealias,
                            double weight, double k_weight, float lr, double* loss);
/* D = A B for 16x16 float matrices (row-major) through one BF16 MFMA, to check
 * the operand layout the kernels assume */
int sh_mfma_probe(const void* A, const void* B, void* D, void* stream);

/* Coarse-grid LES: the FP32 fused solver with f = -div tau_model as its body
 * force, tau from the network at every step (gradient tensor by spectral
 * derivatives, transforms in FP32).  model = NULL is the no-model solve.
 * clip = 1 zeroes tau where it would feed energy to the resolved scales
 * (tau_ij S_ij > 0). */
typedef struct sh_les_s* sh_les_t;
int sh_les_create(sh_les_t* handle, int n, double nu, double K, sh_mlp_t model, int clip, void* stream);
int sh_les_destroy(sh_les_t handle);
/* dealiasing mode of the coarse solver (sh_ns_set_dealias); set it before
 * the initial condition.  Synchronous. */
int sh_les_set_dealias(sh_les_t handle, int mode);
/* initial condition from a physical filtered velocity (float [3][n^3]):
 * transform, zero the Nyquist modes, project onto divergence-free fields,
 * then the start-up step (exec/ns.cpp:132-153) with the closure force */
int sh_les_set_initial_physical(sh_les_t handle, const void* ubar3);
int sh_les_steps(sh_les_t handle, int nsteps);
/* lag = 0 (default): step t is forced by f(V^t), the closure running beside
 * the solver's kernels A-D.  lag = 1: step t is forced by the extrapolation
 * 2 f(V^{t-1}) - f(V^{t-2}) (second order in time, as the AB2 term it joins),
 * so the closure of V^t runs beside the whole solver step t.  Switching on
 * restarts the history at the current state (the next step is forced by
 * f(V)).  Any split of a run into sh_les_steps calls, and graph replays of a
 * call, give the same sequence. */
int sh_les_set_lag(sh_les_t handle, int lag);
/* mixed model: tau = tau_network (after clipping) - 2 (Cs Delta)^2 |S| S, the
 * Smagorinsky stress of the resolved strain rate S = (g + g^T)/2, |S| =
 * sqrt(2 S:S), Delta = 2 pi / n.  cs = 0 (default) is the network alone. */
int sh_les_set_smagorinsky(sh_les_t handle, double cs);
/* tune the solver's and the closure transforms' schedules on this grid
 * (after the initial condition is set; the state is kept).  Synchronous. */
int sh_les_autotune(sh_les_t handle, int reps);
/* velocity in Fourier space, 3 contiguous complex128 fields */
int sh_les_get_velocity(sh_les_t handle, void* out_v3);
/* the model stress of the current velocity, float [6][n^3] physical */
int sh_les_stress(sh_les_t handle, void* out_tau6);
/* the body force -div tau of the current velocity, 3 Fourier complex64 fields */
int sh_les_force(sh_les_t handle, void* out_f3);
/* the force the next step applies with the lagged closure (3 Fourier
 * complex64 fields); with the synchronous closure, the one the last step
 * applied */
int sh_les_pending_force(sh_les_t handle, void* out_f3);
/* network inputs of a physical filtered velocity ubar (float [3][n^3]):
 * X (float [9][ldx]) = d ubar_i / d x_j by spectral derivatives on the n^3
 * grid (Nyquist modes dropped), the same features the LES feeds the model */
int sh_les_features(int n, const void* ubar3, void* X, long ldx, void* stream);

/* ------------------------------------------------------------------ kernel timing
 * When enabled, every launch of the named kernel classes is bracketed with
 * HIP events on its own stream; totals are read after a synchronize. */
int sh_prof_enable(int on);
int sh_prof_reset(void);
/* name: kernel class, e.g. "fft_x", "fft_y", "fft_z", "ns_fourier", "ns_cross_z".
 * Synchronous. */
int sh_prof_query(const char* name, double* total_ms, long* launches);
/* newline-separated list of kernel classes seen so far */
int sh_prof_names(char* buf, size_t len);

#ifdef __cplusplus
}
#endif
#endif
