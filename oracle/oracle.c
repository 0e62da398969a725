/*
 * oracle.c -- CPU restatement of the spectral_upcxx hot path (TEST INFRASTRUCTURE).
 *
 * Checker and timed CPU baseline only; the HIP product never calls it.
 * Every function cites the reference source it restates (paths relative to
 * the reference root).  The 1D engine is a project-written Stockham
 * mixed-radix FFT standing in for FFTW, which is not installed anywhere this
 * code runs; the reference's FFTW calls are plain unnormalised C2C transforms
 * of any length (src/FFT1D.cpp:14-54, 95-101), which is what this computes:
 * radices 4, 2, 3 with dedicated butterflies, and every other prime factor
 * (5, 7, 11, ..., a prime length itself) as a direct DFT stage summed in long
 * double.  Twiddles are evaluated in long double and rounded once.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define OR_PI_L 3.141592653589793238462643383279502884L

/* ------------------------------------------------------------------ PRNG */
uint64_t or_splitmix64(uint64_t seed, uint64_t counter) {
    uint64_t z = seed + (counter + 1) * 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

void or_prng_uniform(uint64_t seed, uint64_t first, size_t n, double* out) {
#pragma omp parallel for schedule(static)
    for (size_t i = 0; i < n; i++)
        out[i] = (double)(or_splitmix64(seed, first + i) >> 11) * 0x1.0p-53;
}

void or_prng_complex(uint64_t seed, int nx, int ny, int nz, double* out) {
    size_t n = (size_t)nx * ny * nz;
#pragma omp parallel for schedule(static)
    for (size_t e = 0; e < 2 * n; e++)
        out[e] = 2.0 * ((double)(or_splitmix64(seed, e) >> 11) * 0x1.0p-53) - 1.0;
}

int or_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* ------------------------------------------------------------------ 1D FFT plans */
typedef struct {
    int n;
    int nstage;
    int radix[40];
    double* w; /* w[2m], w[2m+1] = cos, -sin of 2 pi m / n  (forward twiddles) */
} plan1d;

#define MAX_PLANS 256
static plan1d g_plans[MAX_PLANS];
static int g_nplans = 0;

/* 4s, 2s, 3s, then the remaining prime factors in increasing order (each a
 * direct-DFT stage); any n >= 1 factors */
static int factorize(int n, int* radix) {
    int ns = 0;
    if (n < 1) return -1;
    while (n % 4 == 0) { radix[ns++] = 4; n /= 4; }
    while (n % 2 == 0) { radix[ns++] = 2; n /= 2; }
    while (n % 3 == 0) { radix[ns++] = 3; n /= 3; }
    for (int p = 5; n > 1 && (long)p * p <= n; p += 2)
        while (n % p == 0) { radix[ns++] = p; n /= p; }
    if (n > 1) radix[ns++] = n;
    return ns;
}

static const plan1d* get_plan(int n) {
    const plan1d* found = NULL;
#pragma omp critical(or_plan_cache)
    {
        for (int i = 0; i < g_nplans; i++)
            if (g_plans[i].n == n) found = &g_plans[i];
        if (!found && g_nplans < MAX_PLANS) {
            plan1d* p = &g_plans[g_nplans];
            p->n = n;
            p->nstage = factorize(n, p->radix);
            if (p->nstage >= 0) {
                p->w = (double*)malloc(sizeof(double) * 2 * (size_t)n);
                for (int m = 0; m < n; m++) {
                    long double a = 2.0L * OR_PI_L * (long double)m / (long double)n;
                    p->w[2 * m] = (double)cosl(a);
                    p->w[2 * m + 1] = (double)(-sinl(a));
                }
                g_nplans++;
                found = p;
            }
        }
    }
    return found;
}

/* One Stockham stage of a radix R other than 2, 3, 4 (5, 7 or any larger
 * prime factor): the twiddled inputs, then the direct length-R DFT
 * X[q] = sum_r x[r] exp(sign 2 pi i r q / R) with its constants taken from the
 * length-n table (exp(2 pi i m / R) = w[m n / R]) and the sums in long
 * double. */
static void stage_generic(const plan1d* p, const double* in, double* out, int R, int Ns,
                          double ws) {
    const int n = p->n;
    const int m = n / R;
    const int tws = n / (Ns * R);
    const int step = n / R;
    long double xr[R], xi[R];
    for (int j = 0; j < m; j++) {
        const int k = j % Ns;
        for (int r = 0; r < R; r++) {
            double a = in[2 * (j + r * m)], b = in[2 * (j + r * m) + 1];
            if (Ns > 1 && r > 0) {
                const int t = k * r * tws;
                const double wr = p->w[2 * t], wi = ws * p->w[2 * t + 1];
                const double c = a * wr - b * wi;
                b = a * wi + b * wr;
                a = c;
            }
            xr[r] = a;
            xi[r] = b;
        }
        const int base = (j / Ns) * Ns * R + k;
        for (int q = 0; q < R; q++) {
            long double sr = 0, si = 0;
            for (int r = 0; r < R; r++) {
                const int t = (int)(((long)r * q) % R) * step;
                const long double wr = p->w[2 * t], wi = ws * p->w[2 * t + 1];
                sr += xr[r] * wr - xi[r] * wi;
                si += xr[r] * wi + xi[r] * wr;
            }
            out[2 * (base + q * Ns)] = (double)sr;
            out[2 * (base + q * Ns) + 1] = (double)si;
        }
    }
}

/* Stockham autosort pass over one contiguous line; x is overwritten with the
 * result, y is scratch of the same size. */
static void fft_line(const plan1d* p, double* x, double* y, int dir) {
    const int n = p->n;
    const double ws = (dir < 0) ? 1.0 : -1.0; /* conj twiddles for dir=+1 */
    const double s = (double)dir;             /* sign of the exponent */
    double* in = x;
    double* out = y;
    int Ns = 1;
    for (int st = 0; st < p->nstage; st++) {
        const int R = p->radix[st];
        const int m = n / R;
        const int tws = n / (Ns * R);
        if (R > 4) {
            stage_generic(p, in, out, R, Ns, ws);
            double* t = in; in = out; out = t;
            Ns *= R;
            continue;
        }
        for (int j = 0; j < m; j++) {
            const int k = j % Ns;
            double xr[4], xi[4];
            for (int r = 0; r < R; r++) {
                xr[r] = in[2 * (j + r * m)];
                xi[r] = in[2 * (j + r * m) + 1];
            }
            if (Ns > 1) {
                for (int r = 1; r < R; r++) {
                    int t = k * r * tws;
                    double wr = p->w[2 * t], wi = ws * p->w[2 * t + 1];
                    double a = xr[r] * wr - xi[r] * wi;
                    double b = xr[r] * wi + xi[r] * wr;
                    xr[r] = a;
                    xi[r] = b;
                }
            }
            double Xr[4], Xi[4];
            if (R == 2) {
                Xr[0] = xr[0] + xr[1]; Xi[0] = xi[0] + xi[1];
                Xr[1] = xr[0] - xr[1]; Xi[1] = xi[0] - xi[1];
            } else if (R == 4) {
                double ar = xr[0] + xr[2], ai = xi[0] + xi[2];
                double br = xr[0] - xr[2], bi = xi[0] - xi[2];
                double cr = xr[1] + xr[3], ci = xi[1] + xi[3];
                double er = xr[1] - xr[3], ei = xi[1] - xi[3];
                /* d = e * (s i) */
                double dr = -s * ei, di = s * er;
                Xr[0] = ar + cr; Xi[0] = ai + ci;
                Xr[2] = ar - cr; Xi[2] = ai - ci;
                Xr[1] = br + dr; Xi[1] = bi + di;
                Xr[3] = br - dr; Xi[3] = bi - di;
            } else { /* R == 3 */
                const double h = 0.86602540378443864676; /* sqrt(3)/2 */
                double tr = xr[1] + xr[2], ti = xi[1] + xi[2];
                double mr = xr[0] - 0.5 * tr, mi = xi[0] - 0.5 * ti;
                double er = xr[1] - xr[2], ei = xi[1] - xi[2];
                double dr = -s * h * ei, di = s * h * er;
                Xr[0] = xr[0] + tr; Xi[0] = xi[0] + ti;
                Xr[1] = mr + dr; Xi[1] = mi + di;
                Xr[2] = mr - dr; Xi[2] = mi - di;
            }
            const int base = (j / Ns) * Ns * R + k;
            for (int r = 0; r < R; r++) {
                out[2 * (base + r * Ns)] = Xr[r];
                out[2 * (base + r * Ns) + 1] = Xi[r];
            }
        }
        double* t = in; in = out; out = t;
        Ns *= R;
    }
    if (in != x) memcpy(x, in, sizeof(double) * 2 * (size_t)n);
}

/* FFT1D::transform(ptr, DIR) with the plan of FFT1D::init(N, stride) --
 * src/FFT1D.cpp:14-54 (plans of any N), 95-101 (execution) */
int or_fft1d(double* x, int n, long stride, int dir) {
    const plan1d* p = get_plan(n);
    if (!p) return -1;
    double* a = (double*)malloc(sizeof(double) * 4 * (size_t)n);
    for (int i = 0; i < n; i++) {
        a[2 * i] = x[2 * i * stride];
        a[2 * i + 1] = x[2 * i * stride + 1];
    }
    fft_line(p, a, a + 2 * n, dir);
    for (int i = 0; i < n; i++) {
        x[2 * i * stride] = a[2 * i];
        x[2 * i * stride + 1] = a[2 * i + 1];
    }
    free(a);
    return 0;
}

/* Strided pass: lines of length n at element stride `es`; `nb` adjacent lines
 * (unit distance apart) are gathered together so every cache line is used.
 * 32 columns (512 B per row) keeps power-of-two strides from thrashing the
 * cache sets and TLB (8 columns made a 256^3 transform 16x slower). */
#define OR_BLK 32
static void strided_block(const plan1d* p, double* base, int n, long es, int nb, int dir,
                          double* buf) {
    double* scratch = buf + 2 * (size_t)n * OR_BLK;
    for (int r = 0; r < n; r++) {
        const double* src = base + 2 * (size_t)r * es;
        for (int c = 0; c < nb; c++) {
            buf[2 * ((size_t)c * n + r)] = src[2 * c];
            buf[2 * ((size_t)c * n + r) + 1] = src[2 * c + 1];
        }
    }
    for (int c = 0; c < nb; c++) fft_line(p, buf + 2 * (size_t)c * n, scratch, dir);
    for (int r = 0; r < n; r++) {
        double* dst = base + 2 * (size_t)r * es;
        for (int c = 0; c < nb; c++) {
            dst[2 * c] = buf[2 * ((size_t)c * n + r)];
            dst[2 * c + 1] = buf[2 * ((size_t)c * n + r) + 1];
        }
    }
}

/* Virtual slab ranks of the 3D transform (1 = one rank, OpenMP threads over
 * every loop).  With R > 1 ranks, R threads each own a z-slab of f and a
 * y-slab of tmp and run their phases serially, with the reference's slab
 * puts between them as copies (src/Scalar.cpp:446-503, 582-596). */
static int g_ranks = 1;

void or_set_ranks(int ranks) { g_ranks = ranks < 1 ? 1 : ranks; }
int or_get_ranks(void) { return g_ranks; }

/* Scalar::transform(DIR) at p = R ranks, SLABS variant -- src/Scalar.cpp:400-646,
 * rank r holding planes [r Nz/R, (r+1) Nz/R) of f (its m_local) and the
 * y-slab tmp + r Nx (Ny/R) Nz (its tlocal):
 *   columns (y lines on the local planes, :420-431),
 *   rows (x lines) plane by plane, each plane's row block j sent to rank j at
 *     loc = r (Ny/R) Nx (Nz/R) + k Nx (Ny/R) (:451-474),
 *   barrier, z lines on tlocal at stride Nx (Ny/R) (:530-553),
 *   barrier, transpose back: slab k of tlocal to rank k/(Nz/R) at
 *     dest = r (Ny/R) Nx + (k mod Nz/R) Nx Ny (:570-596). */
#ifdef _OPENMP
static int transform3d_ranks(const plan1d* px, const plan1d* py, const plan1d* pz, double* f,
                             double* tmp, int nx, int ny, int nz, int dir, int R) {
    const long sxy = (long)nx * ny;
    const int nzp = nz / R, nyp = ny / R;
    const int nmax = nx > ny ? (nx > nz ? nx : nz) : (ny > nz ? ny : nz);
#pragma omp parallel num_threads(R)
    {
        const int r = omp_get_thread_num();
        double* buf = (double*)malloc(sizeof(double) * 2 * (size_t)nmax * (OR_BLK + 2));
        double* loc = f + 2 * (size_t)r * nzp * sxy;
        double* tl = tmp + 2 * (size_t)r * nyp * nx * nz;
        for (int k = 0; k < nzp; k++)
            for (int i0 = 0; i0 < nx; i0 += OR_BLK) {
                int nb = nx - i0 < OR_BLK ? nx - i0 : OR_BLK;
                strided_block(py, loc + 2 * ((size_t)k * sxy + i0), ny, nx, nb, dir, buf);
            }
        for (int k = 0; k < nzp; k++) {
            for (int j = 0; j < ny; j++) fft_line(px, loc + 2 * ((size_t)k * sxy + (size_t)j * nx), buf, dir);
            for (int q = 0; q < R; q++) {
                const double* src = loc + 2 * ((size_t)k * sxy + (size_t)q * nx * nyp);
                double* dst = tmp + 2 * ((size_t)q * nyp * nx * nz + (size_t)r * nyp * nx * nzp +
                                         (size_t)k * nx * nyp);
                memcpy(dst, src, sizeof(double) * 2 * (size_t)nx * nyp);
            }
        }
#pragma omp barrier
        for (int j = 0; j < nyp; j++)
            for (int i0 = 0; i0 < nx; i0 += OR_BLK) {
                int nb = nx - i0 < OR_BLK ? nx - i0 : OR_BLK;
                strided_block(pz, tl + 2 * ((size_t)j * nx + i0), nz, (long)nx * nyp, nb, dir, buf);
            }
#pragma omp barrier
        for (int k = 0; k < nz; k++) {
            const double* slab = tl + 2 * (size_t)k * nyp * nx;
            double* dst = f + 2 * ((size_t)(k / nzp) * nzp * sxy + (size_t)r * nyp * nx +
                                   (size_t)(k % nzp) * sxy);
            memcpy(dst, slab, sizeof(double) * 2 * (size_t)nx * nyp);
        }
        free(buf);
    }
    return 0;
}
#endif

/* Scalar::transform(DIR) -- src/Scalar.cpp:400-646.  At one rank the phase
 * order is the reference's: columns (y, :425-431), rows (x, :485-502) with the
 * row copy into tmp (the p=1 branch of the pencil put, :495-496), z on tmp
 * (:547-553), transpose back (slab copies, :582-596), OpenMP threads over the
 * lines of each phase.  With or_set_ranks(R), R | Ny and R | Nz, the slab
 * ranks above. */
int or_transform3d(double* f, double* tmp, int nx, int ny, int nz, int dir) {
    const plan1d* px = get_plan(nx);
    const plan1d* py = get_plan(ny);
    const plan1d* pz = get_plan(nz);
    if (!px || !py || !pz) return -1;
#ifdef _OPENMP
    if (g_ranks > 1 && ny % g_ranks == 0 && nz % g_ranks == 0)
        return transform3d_ranks(px, py, pz, f, tmp, nx, ny, nz, dir, g_ranks);
#endif
    const long sxy = (long)nx * ny;
    const int nmax = nx > ny ? (nx > nz ? nx : nz) : (ny > nz ? ny : nz);
#pragma omp parallel
    {
        double* buf = (double*)malloc(sizeof(double) * 2 * (size_t)nmax * (OR_BLK + 2));
        /* columns: y lines, stride Nx */
#pragma omp for schedule(static) collapse(2)
        for (int k = 0; k < nz; k++)
            for (int i0 = 0; i0 < nx; i0 += OR_BLK) {
                int nb = nx - i0 < OR_BLK ? nx - i0 : OR_BLK;
                strided_block(py, f + 2 * ((size_t)k * sxy + i0), ny, nx, nb, dir, buf);
            }
        /* rows: x lines, contiguous, then copied into tmp */
#pragma omp for schedule(static) collapse(2)
        for (int k = 0; k < nz; k++)
            for (int j = 0; j < ny; j++) {
                double* row = f + 2 * ((size_t)k * sxy + (size_t)j * nx);
                fft_line(px, row, buf, dir);
                memcpy(tmp + 2 * ((size_t)k * sxy + (size_t)j * nx), row,
                       sizeof(double) * 2 * (size_t)nx);
            }
        /* z lines on tmp, stride Nx*Ny */
#pragma omp for schedule(static) collapse(2)
        for (int j = 0; j < ny; j++)
            for (int i0 = 0; i0 < nx; i0 += OR_BLK) {
                int nb = nx - i0 < OR_BLK ? nx - i0 : OR_BLK;
                strided_block(pz, tmp + 2 * ((size_t)j * nx + i0), nz, sxy, nb, dir, buf);
            }
        /* transpose back: slab copies */
#pragma omp for schedule(static)
        for (int k = 0; k < nz; k++)
            memcpy(f + 2 * (size_t)k * sxy, tmp + 2 * (size_t)k * sxy,
                   sizeof(double) * 2 * (size_t)sxy);
        free(buf);
    }
    return 0;
}

/* Scalar::zeroHighModes -- src/Scalar.cpp:336-370 (ZERO on by default) */
void or_zero_high(double* f, int nx, int ny, int nz) {
    const int hx = nx / 2, hy = ny / 2, hz = nz / 2;
#pragma omp parallel for schedule(static)
    for (int k = 0; k < nz; k++) {
        for (int j = 0; j < ny; j++) {
            size_t n = (size_t)hx + (size_t)nx * j + (size_t)nx * ny * k;
            f[2 * n] = 0.0; f[2 * n + 1] = 0.0;
        }
        for (int i = 0; i < nx; i++) {
            size_t n = (size_t)i + (size_t)nx * hy + (size_t)nx * ny * k;
            f[2 * n] = 0.0; f[2 * n + 1] = 0.0;
        }
    }
    double* plane = f + 2 * (size_t)nx * ny * hz;
    memset(plane, 0, sizeof(double) * 2 * (size_t)nx * ny);
}

int or_forward(double* f, double* tmp, int nx, int ny, int nz) {
    int rc = or_transform3d(f, tmp, nx, ny, nz, -1);
    if (rc) return rc;
    or_zero_high(f, nx, ny, nz);
    return 0;
}

int or_inverse(double* f, double* tmp, int nx, int ny, int nz) {
    int rc = or_transform3d(f, tmp, nx, ny, nz, +1);
    if (rc) return rc;
    const size_t n = (size_t)nx * ny * nz;
    const double N = (double)n;
#pragma omp parallel for schedule(static)
    for (size_t i = 0; i < 2 * n; i++) f[i] /= N;
    return 0;
}

/* Scalar::freq -- src/Scalar.cpp:198-206 */
double or_freq(int ind, int n) {
    return (ind <= n / 2) ? (double)ind : -1.0 * (double)n + (double)ind;
}

/* ------------------------------------------------------------------ operators */
#define LOOP3(nx, ny, nz, ...)                                                 \
    _Pragma("omp parallel for schedule(static)") for (int k = 0; k < (nz); k++) \
        for (int j = 0; j < (ny); j++) for (int i = 0; i < (nx); i++) {        \
            const size_t n = (size_t)i + (size_t)(nx) * j + (size_t)(nx) * (ny) * k; \
            const double f0 = or_freq(i, nx), f1 = or_freq(j, ny), f2 = or_freq(k, nz); \
            (void)f0; (void)f1; (void)f2;                                        \
            __VA_ARGS__                                                          \
        }

/* i*f*z with f real: (0, f) * (a, b) = (-f b, f a)  (exact) */
#define IMUL_R(f, a, b) (-(f) * (b))
#define IMUL_I(f, a, b) ((f) * (a))

/* Scalar::gradient -- src/Scalar.cpp:245-276 */
void or_gradient(const double* s, double* vx, double* vy, double* vz, int nx, int ny, int nz) {
    LOOP3(nx, ny, nz, {
        double a = s[2 * n], b = s[2 * n + 1];
        vx[2 * n] = IMUL_R(f0, a, b); vx[2 * n + 1] = IMUL_I(f0, a, b);
        vy[2 * n] = IMUL_R(f1, a, b); vy[2 * n + 1] = IMUL_I(f1, a, b);
        vz[2 * n] = IMUL_R(f2, a, b); vz[2 * n + 1] = IMUL_I(f2, a, b);
    })
}

/* Scalar::laplacian -- src/Scalar.cpp:278-306 */
void or_laplacian(const double* s, double* out, int nx, int ny, int nz) {
    LOOP3(nx, ny, nz, {
        double q = -(f0 * f0 + f1 * f1 + f2 * f2);
        out[2 * n] = q * s[2 * n];
        out[2 * n + 1] = q * s[2 * n + 1];
    })
}

/* Scalar::laplacian_inverse(a, b) -- src/Scalar.cpp:308-334 */
void or_laplacian_inverse(double* s, double a, double b, int nx, int ny, int nz) {
    LOOP3(nx, ny, nz, {
        double sum = 0;
        sum += f0 * f0; sum += f1 * f1; sum += f2 * f2;
        if (sum != 0) {
            double d = a - b * sum;
            s[2 * n] /= d;
            s[2 * n + 1] /= d;
        }
    })
}

/* Vector::divergence -- src/Vector.cpp:124-153 */
void or_divergence(const double* vx, const double* vy, const double* vz, double* out,
                   int nx, int ny, int nz) {
    LOOP3(nx, ny, nz, {
        double re = 0, im = 0;
        re += IMUL_R(f0, vx[2 * n], vx[2 * n + 1]); im += IMUL_I(f0, vx[2 * n], vx[2 * n + 1]);
        re += IMUL_R(f1, vy[2 * n], vy[2 * n + 1]); im += IMUL_I(f1, vy[2 * n], vy[2 * n + 1]);
        re += IMUL_R(f2, vz[2 * n], vz[2 * n + 1]); im += IMUL_I(f2, vz[2 * n], vz[2 * n + 1]);
        out[2 * n] = re; out[2 * n + 1] = im;
    })
}

/* Vector::curl -- src/Vector.cpp:92-122 */
void or_curl(const double* vx, const double* vy, const double* vz,
             double* ox, double* oy, double* oz, int nx, int ny, int nz) {
    LOOP3(nx, ny, nz, {
        double xr = vx[2 * n], xi = vx[2 * n + 1];
        double yr = vy[2 * n], yi = vy[2 * n + 1];
        double zr = vz[2 * n], zi = vz[2 * n + 1];
        ox[2 * n] = IMUL_R(f1, zr, zi) - IMUL_R(f2, yr, yi);
        ox[2 * n + 1] = IMUL_I(f1, zr, zi) - IMUL_I(f2, yr, yi);
        oy[2 * n] = IMUL_R(f2, xr, xi) - IMUL_R(f0, zr, zi);
        oy[2 * n + 1] = IMUL_I(f2, xr, xi) - IMUL_I(f0, zr, zi);
        oz[2 * n] = IMUL_R(f0, yr, yi) - IMUL_R(f1, xr, xi);
        oz[2 * n + 1] = IMUL_I(f0, yr, yi) - IMUL_I(f1, xr, xi);
    })
}

/* Vector::laplacian -- src/Vector.cpp:155-181 */
void or_vlaplacian(const double* vx, const double* vy, const double* vz,
                   double* ox, double* oy, double* oz, int nx, int ny, int nz) {
    LOOP3(nx, ny, nz, {
        double q = -(f0 * f0 + f1 * f1 + f2 * f2);
        ox[2 * n] = q * vx[2 * n]; ox[2 * n + 1] = q * vx[2 * n + 1];
        oy[2 * n] = q * vy[2 * n]; oy[2 * n + 1] = q * vy[2 * n + 1];
        oz[2 * n] = q * vz[2 * n]; oz[2 * n + 1] = q * vz[2 * n + 1];
    })
}

/* Vector::cross -- src/Vector.cpp:66-90: inverse copies of both operands,
 * complex pointwise u x v, forward (with Nyquist zeroing) of the result. */
int or_cross(const double* a, const double* b, double* out, double* work, int nx, int ny, int nz) {
    const size_t n = (size_t)nx * ny * nz;
    const size_t F = 2 * n;
    double* u = work;          /* 3 fields */
    double* v = work + 3 * F;  /* 3 fields */
    double* tmp = work + 6 * F;
    memcpy(u, a, sizeof(double) * 3 * F);
    memcpy(v, b, sizeof(double) * 3 * F);
    for (int d = 0; d < 3; d++) {
        int rc = or_inverse(u + d * F, tmp, nx, ny, nz);
        if (rc) return rc;
        rc = or_inverse(v + d * F, tmp, nx, ny, nz);
        if (rc) return rc;
    }
#pragma omp parallel for schedule(static)
    for (size_t i = 0; i < n; i++) {
        double u0r = u[2 * i], u0i = u[2 * i + 1];
        double u1r = u[F + 2 * i], u1i = u[F + 2 * i + 1];
        double u2r = u[2 * F + 2 * i], u2i = u[2 * F + 2 * i + 1];
        double v0r = v[2 * i], v0i = v[2 * i + 1];
        double v1r = v[F + 2 * i], v1i = v[F + 2 * i + 1];
        double v2r = v[2 * F + 2 * i], v2i = v[2 * F + 2 * i + 1];
        /* ret[0] = u1 v2 - u2 v1 ; ret[1] = u2 v0 - u0 v2 ; ret[2] = u0 v1 - u1 v0 */
        out[2 * i] = (u1r * v2r - u1i * v2i) - (u2r * v1r - u2i * v1i);
        out[2 * i + 1] = (u1r * v2i + u1i * v2r) - (u2r * v1i + u2i * v1r);
        out[F + 2 * i] = (u2r * v0r - u2i * v0i) - (u0r * v2r - u0i * v2i);
        out[F + 2 * i + 1] = (u2r * v0i + u2i * v0r) - (u0r * v2i + u0i * v2r);
        out[2 * F + 2 * i] = (u0r * v1r - u0i * v1i) - (u1r * v0r - u1i * v0i);
        out[2 * F + 2 * i + 1] = (u0r * v1i + u0i * v1r) - (u1r * v0i + u1i * v0r);
    }
    for (int d = 0; d < 3; d++) {
        int rc = or_forward(out + d * F, tmp, nx, ny, nz);
        if (rc) return rc;
    }
    return 0;
}

/* Scalar::average -- src/Scalar.cpp:15-22: sum of abs(m_local[i]) over the
 * local data divided by localSize() (at one rank: every element) */
double or_average(const double* f, size_t n) {
    double acc = 0;
#pragma omp parallel for schedule(static) reduction(+ : acc)
    for (size_t i = 0; i < n; i++) acc += hypot(f[2 * i], f[2 * i + 1]);
    return acc / (double)n;
}

/* ------------------------------------------------------------------ elementwise helpers */
static void v_add(double* out, const double* x, const double* y, size_t m) {
#pragma omp parallel for schedule(static)
    for (size_t i = 0; i < m; i++) out[i] = x[i] + y[i];
}
static void v_sub(double* out, const double* x, const double* y, size_t m) {
#pragma omp parallel for schedule(static)
    for (size_t i = 0; i < m; i++) out[i] = x[i] - y[i];
}
static void v_scale(double* out, double a, const double* x, size_t m) {
#pragma omp parallel for schedule(static)
    for (size_t i = 0; i < m; i++) out[i] = a * x[i];
}
static void v_copy(double* out, const double* x, size_t m) {
#pragma omp parallel for schedule(static)
    for (size_t i = 0; i < m; i++) out[i] = x[i];
}

/* ------------------------------------------------------------------ NS driver */
struct or_ns {
    int nx, ny, nz;
    size_t F; /* doubles per field */
    double nu, K;
    int nonlinear;
    int scheme; /* 0: exec/ns.cpp, 1: exec/ns2.cpp */
    double div_avg;
    /* vectors: 3 contiguous fields each */
    double *V0, *V1, *V, *omega0, *omega1, *omega, *cross0, *cross1, *F3, *lap1, *S, *Gg, *Vh;
    double *Pi0, *G, *div;
    double* work; /* 7 fields for or_cross */
};

static double* falloc(size_t m) {
    double* p = (double*)malloc(sizeof(double) * m);
    if (p) {
#pragma omp parallel for schedule(static)
        for (size_t i = 0; i < m; i++) p[i] = 0.0;
    }
    return p;
}

or_ns* or_ns_create(int nx, int ny, int nz, double nu, double K) {
    or_ns* s = (or_ns*)calloc(1, sizeof(or_ns));
    s->nx = nx; s->ny = ny; s->nz = nz;
    s->F = 2 * (size_t)nx * ny * nz;
    s->nu = nu; s->K = K; s->nonlinear = 1; s->scheme = 0;
    size_t V = 3 * s->F;
    s->V0 = falloc(V); s->V1 = falloc(V); s->V = falloc(V);
    s->omega0 = falloc(V); s->omega1 = falloc(V); s->omega = falloc(V);
    s->cross0 = falloc(V); s->cross1 = falloc(V); s->F3 = falloc(V);
    s->lap1 = falloc(V); s->S = falloc(V); s->Gg = falloc(V); s->Vh = falloc(V);
    s->Pi0 = falloc(s->F); s->G = falloc(s->F); s->div = falloc(s->F);
    s->work = falloc(7 * s->F);
    return s;
}

void or_ns_destroy(or_ns* s) {
    if (!s) return;
    double* ptrs[] = {s->V0, s->V1, s->V, s->omega0, s->omega1, s->omega, s->cross0, s->cross1,
                      s->F3, s->lap1, s->S, s->Gg, s->Vh, s->Pi0, s->G, s->div, s->work};
    for (size_t i = 0; i < sizeof(ptrs) / sizeof(ptrs[0]); i++) free(ptrs[i]);
    free(s);
}

void or_ns_set_nonlinear(or_ns* s, int on) { s->nonlinear = on; }

/* 0: exec/ns.cpp (Crank-Nicolson viscous term), 1: exec/ns2.cpp (backward
 * Euler viscous term, exec/ns2.cpp:171-189); the start-up is the same
 * (exec/ns2.cpp:133-137 = exec/ns.cpp:132-153) */
void or_ns_set_scheme(or_ns* s, int scheme) { s->scheme = scheme; }

#define FX(p) (p)
#define FY(p) ((p) + s->F)
#define FZ(p) ((p) + 2 * s->F)

static int ns_cross(or_ns* s, const double* a, const double* b, double* out) {
    if (!s->nonlinear) {
        memset(out, 0, sizeof(double) * 3 * s->F);
        return 0;
    }
    return or_cross(a, b, out, s->work, s->nx, s->ny, s->nz);
}

static void ns_vlap(or_ns* s, const double* v, double* out) {
    or_vlaplacian(FX(v), FY(v), FZ(v), FX(out), FY(out), FZ(out), s->nx, s->ny, s->nz);
}

int or_ns_set_initial(or_ns* s, const double* V0) {
    v_copy(s->V0, V0, 3 * s->F);
    or_curl(FX(s->V0), FY(s->V0), FZ(s->V0), FX(s->omega0), FY(s->omega0), FZ(s->omega0),
            s->nx, s->ny, s->nz);
    return 0;
}

/* exec/ns.cpp:113-135, exec/ns2.cpp:113-137: the drivers start from the
 * vorticity they built V0 from, not from curl(V0) */
void or_ns_set_vorticity(or_ns* s, const double* omega0) { v_copy(s->omega0, omega0, 3 * s->F); }

/* exec/ns.cpp:132-153 */
int or_ns_startup(or_ns* s) {
    const size_t V = 3 * s->F;
    const int nx = s->nx, ny = s->ny, nz = s->nz;
    int rc;
    /* Pi0 = (V0.cross(omega0) + nu*V0.laplacian()).divergence(); Pi0.laplacian_inverse() */
    if ((rc = ns_cross(s, s->V0, s->omega0, s->cross1))) return rc;
    ns_vlap(s, s->V0, s->lap1);
    v_scale(s->lap1, s->nu, s->lap1, V);
    v_add(s->S, s->cross1, s->lap1, V);
    or_divergence(FX(s->S), FY(s->S), FZ(s->S), s->Pi0, nx, ny, nz);
    or_laplacian_inverse(s->Pi0, 0.0, 1.0, nx, ny, nz);
    /* V1 = V0 + K*(V0.cross(omega0) + nu*V0.laplacian() - Pi0.gradient()) */
    or_gradient(s->Pi0, FX(s->Gg), FY(s->Gg), FZ(s->Gg), nx, ny, nz);
    v_sub(s->S, s->S, s->Gg, V); /* (cross + nu lap) - grad Pi0 */
    v_scale(s->S, s->K, s->S, V);
    v_add(s->V1, s->V0, s->S, V);
    /* omega1 = V1.curl() */
    or_curl(FX(s->V1), FY(s->V1), FZ(s->V1), FX(s->omega1), FY(s->omega1), FZ(s->omega1),
            nx, ny, nz);
    /* cross0 = V0.cross(omega0) */
    if ((rc = ns_cross(s, s->V0, s->omega0, s->cross0))) return rc;
    return 0;
}

/* exec/ns.cpp:167-217, one iteration per loop trip */
int or_ns_steps(or_ns* s, int nsteps) {
    const size_t V = 3 * s->F;
    const int nx = s->nx, ny = s->ny, nz = s->nz;
    const double K = s->K, nu = s->nu;
    for (int t = 0; t < nsteps; t++) {
        int rc;
        /* cross1 = V1.cross(omega1) */
        if ((rc = ns_cross(s, s->V1, s->omega1, s->cross1))) return rc;
        /* F = K/2*(3.*cross1 - cross0) */
        v_scale(s->F3, 3.0, s->cross1, V);
        v_sub(s->F3, s->F3, s->cross0, V);
        v_scale(s->F3, K / 2, s->F3, V);
        /* lap1 = nu*K/2*V1.laplacian()  (exec/ns2.cpp:178: nu*K*V1.laplacian()) */
        ns_vlap(s, s->V1, s->lap1);
        v_scale(s->lap1, s->scheme ? nu * K : nu * K / 2, s->lap1, V);
        /* G = (V1 + lap1 + F).divergence(); G.laplacian_inverse() */
        v_add(s->S, s->V1, s->lap1, V);
        v_add(s->S, s->S, s->F3, V);
        or_divergence(FX(s->S), FY(s->S), FZ(s->S), s->G, nx, ny, nz);
        or_laplacian_inverse(s->G, 0.0, 1.0, nx, ny, nz);
        /* Vhalf = V1 + F - G.gradient() */
        or_gradient(s->G, FX(s->Gg), FY(s->Gg), FZ(s->Gg), nx, ny, nz);
        v_add(s->Vh, s->V1, s->F3, V);
        v_sub(s->Vh, s->Vh, s->Gg, V);
        if (s->scheme == 0) {
            /* V = Vhalf + lap1; V[d].laplacian_inverse(1, -nu*K/2) */
            v_add(s->V, s->Vh, s->lap1, V);
            for (int d = 0; d < 3; d++)
                or_laplacian_inverse(s->V + d * s->F, 1.0, -nu * K / 2, nx, ny, nz);
        } else {
            /* exec/ns2.cpp:185-189: V = Vhalf; V[d].laplacian_inverse(1, -nu*K) */
            v_copy(s->V, s->Vh, V);
            for (int d = 0; d < 3; d++)
                or_laplacian_inverse(s->V + d * s->F, 1.0, -nu * K, nx, ny, nz);
        }
        /* omega = V.curl(); div = V.divergence(); div.average() */
        or_curl(FX(s->V), FY(s->V), FZ(s->V), FX(s->omega), FY(s->omega), FZ(s->omega),
                nx, ny, nz);
        or_divergence(FX(s->V), FY(s->V), FZ(s->V), s->div, nx, ny, nz);
        s->div_avg = or_average(s->div, s->F / 2);
        /* histories */
        v_copy(s->V0, s->V1, V);
        v_copy(s->V1, s->V, V);
        v_copy(s->omega0, s->omega1, V);
        v_copy(s->omega1, s->omega, V);
        v_copy(s->cross0, s->cross1, V);
    }
    return 0;
}

/* Timing helper: enter the loop state directly (V1 = V0, omega1 = omega0,
 * cross0 = 0) without the start-up's extra transforms.  The per-step work of
 * or_ns_steps does not depend on the values, so a timed window after this is
 * the same work as after or_ns_startup. */
int or_ns_prime_for_timing(or_ns* s) {
    const size_t V = 3 * s->F;
    v_copy(s->V1, s->V0, V);
    v_copy(s->omega1, s->omega0, V);
    memset(s->cross0, 0, sizeof(double) * V);
    return 0;
}

void or_ns_get_velocity(const or_ns* s, double* out) { memcpy(out, s->V1, sizeof(double) * 3 * s->F); }
void or_ns_get_vorticity(const or_ns* s, double* out) { memcpy(out, s->omega1, sizeof(double) * 3 * s->F); }
double or_ns_div_average(const or_ns* s) { return s->div_avg; }

/* ------------------------------------------------------------------ initial conditions */
/* Taylor-Green u = (sin x cos y cos z, -cos x sin y cos z, 0) on x_i = 2 pi i / N
 * (grid as exec/ns.cpp:93-103), then forward() per component. */
int or_ic_taylor_green(double* V, double* tmp, int nx, int ny, int nz) {
    const size_t F = 2 * (size_t)nx * ny * nz;
    const double hx = 2. * M_PI / nx, hy = 2. * M_PI / ny, hz = 2. * M_PI / nz;
#pragma omp parallel for schedule(static)
    for (int k = 0; k < nz; k++)
        for (int j = 0; j < ny; j++)
            for (int i = 0; i < nx; i++) {
                size_t n = (size_t)i + (size_t)nx * j + (size_t)nx * ny * k;
                double x = hx * i, y = hy * j, z = hz * k;
                V[2 * n] = sin(x) * cos(y) * cos(z);
                V[2 * n + 1] = 0;
                V[F + 2 * n] = -cos(x) * sin(y) * cos(z);
                V[F + 2 * n + 1] = 0;
                V[2 * F + 2 * n] = 0;
                V[2 * F + 2 * n + 1] = 0;
            }
    for (int d = 0; d < 3; d++) {
        int rc = or_forward(V + d * F, tmp, nx, ny, nz);
        if (rc) return rc;
    }
    return 0;
}

/* Random solenoidal field, E(k) ~ k^4 exp(-2 (k/k0)^2): per-mode amplitude
 * sqrt(E/(4 pi k^2)), phases 2 pi U(seed, 3m+d), projection I - k k^T/k^2,
 * Hermitian symmetrisation, k=0 and Nyquist planes zeroed, scaled to u_rms. */
void or_ic_solenoidal(double* V, int nx, int ny, int nz, uint64_t seed, double k0, double urms) {
    const size_t n = (size_t)nx * ny * nz;
    const size_t F = 2 * n;
    double* A = (double*)malloc(sizeof(double) * 3 * F);
#pragma omp parallel for schedule(static)
    for (int k = 0; k < nz; k++)
        for (int j = 0; j < ny; j++)
            for (int i = 0; i < nx; i++) {
                size_t m = (size_t)i + (size_t)nx * j + (size_t)nx * ny * k;
                double kx = or_freq(i, nx), ky = or_freq(j, ny), kz = or_freq(k, nz);
                double q = kx * kx + ky * ky + kz * kz;
                double ur[3], ui[3];
                if (q > 0) {
                    double km = sqrt(q);
                    double E = pow(km, 4) * exp(-2.0 * (km / k0) * (km / k0));
                    double amp = sqrt(E / (4.0 * M_PI * q));
                    for (int d = 0; d < 3; d++) {
                        double ph = 2.0 * M_PI *
                                    ((double)(or_splitmix64(seed, 3 * m + d) >> 11) * 0x1.0p-53);
                        ur[d] = amp * cos(ph);
                        ui[d] = amp * sin(ph);
                    }
                    double kr = kx * ur[0] + ky * ur[1] + kz * ur[2];
                    double ki = kx * ui[0] + ky * ui[1] + kz * ui[2];
                    double inv = 1.0 / q;
                    double kk[3] = {kx, ky, kz};
                    for (int d = 0; d < 3; d++) {
                        ur[d] = ur[d] - kk[d] * (kr * inv);
                        ui[d] = ui[d] - kk[d] * (ki * inv);
                    }
                } else {
                    for (int d = 0; d < 3; d++) ur[d] = ui[d] = 0;
                }
                for (int d = 0; d < 3; d++) {
                    A[d * F + 2 * m] = ur[d];
                    A[d * F + 2 * m + 1] = ui[d];
                }
            }
    /* Hermitian symmetrisation and zeroing */
#pragma omp parallel for schedule(static)
    for (int k = 0; k < nz; k++)
        for (int j = 0; j < ny; j++)
            for (int i = 0; i < nx; i++) {
                size_t m = (size_t)i + (size_t)nx * j + (size_t)nx * ny * k;
                int in = (nx - i) % nx, jn = (ny - j) % ny, kn = (nz - k) % nz;
                size_t mn = (size_t)in + (size_t)nx * jn + (size_t)nx * ny * kn;
                int zero = (i == nx / 2) || (j == ny / 2) || (k == nz / 2) || m == 0;
                for (int d = 0; d < 3; d++) {
                    double re = 0.5 * (A[d * F + 2 * m] + A[d * F + 2 * mn]);
                    double im = 0.5 * (A[d * F + 2 * m + 1] - A[d * F + 2 * mn + 1]);
                    V[d * F + 2 * m] = zero ? 0.0 : re;
                    V[d * F + 2 * m + 1] = zero ? 0.0 : im;
                }
            }
    free(A);
    double e2 = 0;
#pragma omp parallel for schedule(static) reduction(+ : e2)
    for (size_t i = 0; i < 3 * F; i++) e2 += V[i] * V[i];
    e2 /= (double)n * (double)n;
    double sc = urms / sqrt(e2 / 3.0);
#pragma omp parallel for schedule(static)
    for (size_t i = 0; i < 3 * F; i++) V[i] *= sc;
}

double or_energy(const double* V, int nx, int ny, int nz) {
    const size_t n = (size_t)nx * ny * nz;
    double e2 = 0;
#pragma omp parallel for schedule(static) reduction(+ : e2)
    for (size_t i = 0; i < 6 * n; i++) e2 += V[i] * V[i];
    return 0.5 * e2 / ((double)n * (double)n);
}

/* Dissipation rate 2 nu sum_modes |k|^2 e(k) of a Fourier velocity, e(k) the
 * mode's kinetic energy 0.5 |u(k)|^2 / N^2 (numpy ifftn scaling, as
 * or_energy); for a solenoidal field it equals 2 nu times the enstrophy. */
double or_dissipation(const double* V, int nx, int ny, int nz, double nu) {
    const size_t n = (size_t)nx * ny * nz;
    const size_t F = 2 * n;
    double acc = 0;
#pragma omp parallel for schedule(static) reduction(+ : acc)
    for (int k = 0; k < nz; k++)
        for (int j = 0; j < ny; j++)
            for (int i = 0; i < nx; i++) {
                size_t m = (size_t)i + (size_t)nx * j + (size_t)nx * ny * k;
                double kx = or_freq(i, nx), ky = or_freq(j, ny), kz = or_freq(k, nz);
                double e = 0;
                for (int d = 0; d < 3; d++)
                    e += V[d * F + 2 * m] * V[d * F + 2 * m] + V[d * F + 2 * m + 1] * V[d * F + 2 * m + 1];
                acc += (kx * kx + ky * ky + kz * kz) * 0.5 * e;
            }
    return 2.0 * nu * acc / ((double)n * (double)n);
}

/* Shell spectrum E(b) = sum of the mode energies e(k) (as or_dissipation)
 * over the modes with round(|k|) = b, b < nbins (modes beyond are dropped). */
void or_spectrum(const double* V, int nx, int ny, int nz, double* Ek, int nbins) {
    const size_t n = (size_t)nx * ny * nz;
    const size_t F = 2 * n;
    for (int b = 0; b < nbins; b++) Ek[b] = 0;
    for (int k = 0; k < nz; k++)
        for (int j = 0; j < ny; j++)
            for (int i = 0; i < nx; i++) {
                size_t m = (size_t)i + (size_t)nx * j + (size_t)nx * ny * k;
                double kx = or_freq(i, nx), ky = or_freq(j, ny), kz = or_freq(k, nz);
                int b = (int)floor(sqrt(kx * kx + ky * ky + kz * kz) + 0.5);
                if (b >= nbins) continue;
                double e = 0;
                for (int d = 0; d < 3; d++)
                    e += V[d * F + 2 * m] * V[d * F + 2 * m] + V[d * F + 2 * m + 1] * V[d * F + 2 * m + 1];
                Ek[b] += 0.5 * e / ((double)n * (double)n);
            }
}

/* Enstrophy 0.5 <|curl u|^2> and the divergence norm sqrt(<|div u|^2>) of a
 * Fourier velocity (numpy ifftn scaling, as or_energy); the curl and the
 * divergence are those of src/Vector.cpp:92-138 (i k x V, i k . V). */
void or_vorticity_stats(const double* V, int nx, int ny, int nz, double* enstrophy, double* divnorm) {
    const size_t n = (size_t)nx * ny * nz;
    const size_t F = 2 * n;
    double w2 = 0, d2 = 0;
#pragma omp parallel for schedule(static) reduction(+ : w2, d2)
    for (int k = 0; k < nz; k++)
        for (int j = 0; j < ny; j++)
            for (int i = 0; i < nx; i++) {
                size_t m = (size_t)i + (size_t)nx * j + (size_t)nx * ny * k;
                double kk[3] = {or_freq(i, nx), or_freq(j, ny), or_freq(k, nz)};
                double re[3], im[3];
                for (int d = 0; d < 3; d++) {
                    re[d] = V[d * F + 2 * m];
                    im[d] = V[d * F + 2 * m + 1];
                }
                /* |k x u|^2 and |k . u|^2 (the factor i has modulus 1) */
                for (int d = 0; d < 3; d++) {
                    int a = (d + 1) % 3, b = (d + 2) % 3;
                    double cr = kk[a] * re[b] - kk[b] * re[a];
                    double ci = kk[a] * im[b] - kk[b] * im[a];
                    w2 += cr * cr + ci * ci;
                }
                double dr = kk[0] * re[0] + kk[1] * re[1] + kk[2] * re[2];
                double di = kk[0] * im[0] + kk[1] * im[1] + kk[2] * im[2];
                d2 += dr * dr + di * di;
            }
    const double nn = (double)n * (double)n;
    if (enstrophy) *enstrophy = 0.5 * w2 / nn;
    if (divnorm) *divnorm = sqrt(d2 / nn);
}

/* ------------------------------------------------------------------ Gaussian vortices */
static double gaussian2(double x, double y, double xc, double yc, double width) {
    return exp(-pow(x - xc, 2) / width) * exp(-pow(y - yc, 2) / width);
}

/* exec/ns.cpp:84-131 (set 0, three vortices) and exec/ns2.cpp:84-130 (set 1,
 * two vortices): omega0 = (0, 0, w_z), forward() per component; psi =
 * -w_z/(kx^2+ky^2) where (kx, ky) != 0 and 0 elsewhere; V0 = (-grad_y psi,
 * grad_x psi, 0).  tmp has one field. */
int or_ic_vortices(double* V0, double* omega0, double* tmp, int nx, int ny, int nz, int set) {
    const size_t F = 2 * (size_t)nx * ny * nz;
    const double width = M_PI / 8, mag = 50, dist = M_PI / 2;
    const double hx = 2. * M_PI / nx, hy = 2. * M_PI / ny;
    for (size_t i = 0; i < 3 * F; i++) omega0[i] = V0[i] = 0;
    double* wz = omega0 + 2 * F;
#pragma omp parallel for schedule(static)
    for (int k = 0; k < nz; k++)
        for (int j = 0; j < ny; j++)
            for (int i = 0; i < nx; i++) {
                size_t m = (size_t)i + (size_t)nx * j + (size_t)nx * ny * k;
                double x = hx * i, y = hy * j;
                double w;
                if (set == 0)
                    w = mag * gaussian2(x, y, M_PI - dist, M_PI - dist, width) +
                        mag * gaussian2(x, y, M_PI + dist, M_PI - dist, width) +
                        mag * gaussian2(x, y, M_PI, M_PI + dist, width);
                else
                    w = mag * gaussian2(x, y, M_PI - dist, M_PI, width) +
                        mag * gaussian2(x, y, M_PI + dist, M_PI, width);
                wz[2 * m] = w;
            }
    for (int d = 0; d < 3; d++) {
        int rc = or_forward(omega0 + d * F, tmp, nx, ny, nz);
        if (rc) return rc;
    }
    double* psi = tmp;
#pragma omp parallel for schedule(static)
    for (int k = 0; k < nz; k++)
        for (int j = 0; j < ny; j++)
            for (int i = 0; i < nx; i++) {
                size_t m = (size_t)i + (size_t)nx * j + (size_t)nx * ny * k;
                double kx = or_freq(i, nx), ky = or_freq(j, ny);
                if (kx != 0 || ky != 0) {
                    double c = -1. / (kx * kx + ky * ky);
                    psi[2 * m] = c * wz[2 * m];
                    psi[2 * m + 1] = c * wz[2 * m + 1];
                } else {
                    psi[2 * m] = psi[2 * m + 1] = 0;
                }
            }
    /* grad = psi.gradient(); V0[0] = -1.*grad[1]; V0[1] = grad[0]
     * (gradient written as x -> field 0, y -> field 2, z -> field 1, then moved) */
    or_gradient(psi, V0, V0 + 2 * F, V0 + F, nx, ny, nz);
    for (size_t i = 0; i < F; i++) {
        double gy = V0[2 * F + i];
        V0[F + i] = V0[i];
        V0[i] = -1. * gy;
        V0[2 * F + i] = 0;
    }
    return 0;
}

/* ------------------------------------------------------------------ marker particles */
/* Particle::Particle (src/Particle.cpp:5-10) places each particle uniformly in
 * [0, 2 pi]^3 with rand(); here coordinate d of particle p is 2 pi U(seed,
 * 3p + d) (splitmix64, 53-bit), reproducible on every platform. */
void or_particles_seed(double* loc, long np, uint64_t seed) {
    for (long p = 0; p < 3 * np; p++) loc[p] = (double)(or_splitmix64(seed, (uint64_t)p) >> 11) * 0x1.0p-53 * 2 * M_PI;
}

/* Particle::move (src/Particle.cpp:12-46) for np particles, loc and vel [np][3],
 * U the physical velocity (3 complex fields, real parts read).  Written in
 * the reference's operation order with contraction off, so that a device
 * implementation doing the same IEEE operations matches bit for bit.  Cell
 * indices are clamped to [0, dims-2]: the upper clamp is the reference's,
 * the lower one only guards positions outside [0, 2 pi), which the
 * reference's wrap never produces for |v K| < 2 pi. */
__attribute__((optimize("fp-contract=off")))
void or_particles_move(double* loc, double* vel, double* energy, long np, const double* U,
                       int nx, int ny, int nz, double K) {
    const size_t F = 2 * (size_t)nx * ny * nz;
    const long dims[3] = {nx, ny, nz};
    for (long p = 0; p < np; p++) {
        double* x = loc + 3 * p;
        double* v = vel + 3 * p;
        long ind[3];
        double rd[3];
        for (int d = 0; d < 3; d++) {
            double h = 2 * M_PI / dims[d];
            double q = x[d] / h;
            long i = q >= 0 ? (q < (double)dims[d] ? (long)q : dims[d]) : 0;
            if (i + 1 >= dims[d]) i = dims[d] - 2;
            ind[d] = i;
            rd[d] = (x[d] - (double)i * h) / h;
        }
#define UAT(d, a, b, c) U[(d) * F + 2 * ((size_t)(ind[0] + (a)) + (size_t)nx * (ind[1] + (b)) + (size_t)nx * ny * (ind[2] + (c)))]
        for (int d = 0; d < 3; d++) {
            v[d] = UAT(d, 0, 0, 0) * (1 - rd[0]) * (1 - rd[1]) * (1 - rd[2]) +
                   UAT(d, 1, 0, 0) * rd[0] * (1 - rd[1]) * (1 - rd[2]) +
                   UAT(d, 1, 1, 0) * rd[0] * rd[1] * (1 - rd[2]) +
                   UAT(d, 0, 1, 0) * (1 - rd[0]) * rd[1] * (1 - rd[2]) +
                   UAT(d, 0, 0, 1) * (1 - rd[0]) * (1 - rd[1]) * (rd[2]) +
                   UAT(d, 1, 0, 1) * rd[0] * (1 - rd[1]) * (rd[2]) +
                   UAT(d, 1, 1, 1) * rd[0] * rd[1] * (rd[2]) +
                   UAT(d, 0, 1, 1) * (1 - rd[0]) * rd[1] * (rd[2]);
        }
#undef UAT
        for (int d = 0; d < 3; d++) {
            x[d] += v[d] * K;
            if (x[d] >= 2 * M_PI)
                x[d] = x[d] - 2 * M_PI;
            else if (x[d] <= 0)
                x[d] = 2 * M_PI + x[d];
        }
        if (energy) {
            double e = 0;
            for (int d = 0; d < 3; d++) e += v[d] * v[d];
            energy[p] = .5 * e;
        }
    }
}
