/*
 * tdse_oracle.c -- CPU FP64 restatement of the reference split-operator path.
 *
 * TEST INFRASTRUCTURE ONLY: the parity checker for the HIP propagator.  Only
 * tests/, __graft_entry__.smoke() and the cpu_baseline leg of bench.py load
 * this library; the product path never does.
 *
 * Built with -O2 -ffp-contract=off and no fast-math, the flag contract of the
 * reference core (CMakeLists.txt:59-112), so that every expression below
 * rounds exactly like its reference counterpart.  Operand order inside each
 * expression follows the reference text: IEEE addition and multiplication are
 * commutative but not associative, so the grouping is what matters.
 *
 * Complex numbers are interleaved (re, im) doubles.  A complex product is the
 * textbook formula (ac - bd, ad + bc): the reference builds with
 * -fcx-limited-range, which makes std::complex<double>::operator* exactly that
 * (CMakeLists.txt:83-95).  std::norm(z) in libstdc++ is re*re + im*im.
 *
 * Pinning: fft/ifft/step/builders are compared bit for bit against the
 * reference core compiled from its own sources (oracle/Makefile, target
 * "ref") and against the closed forms of the reference tests
 * (tests/fft_test.cpp, tests/fft3_test.cpp, tests/propagator_test.cpp,
 * tests/potential_test.cpp).  The mixed-radix FFT has no reference
 * counterpart; it is pinned by the brute-force DFT and numpy.fft.
 */
#include "tdse_oracle.h"

#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define ORC_PI 3.14159265358979323846

static inline void cmul(double ar, double ai, double br, double bi, double* re, double* im) {
    *re = ar * br - ai * bi;
    *im = ar * bi + ai * br;
}

/* ---- grid / spectral ---------------------------------------------------- */

/* core/src/grid.ixx:15-17: spacing = (xmax - xmin) / n, coord = xmin + i*h */
double orc_spacing(double xmin, double xmax, int n) { return (xmax - xmin) / n; }

static inline double coord(double xmin, double xmax, int n, int i) {
    return xmin + i * orc_spacing(xmin, xmax, n);
}

/* core/src/spectral.ixx:14-23: fftfreq order, DC at n = 1 */
void orc_wavenumbers(int n, double xmin, double xmax, double* k) {
    const double dk = 2.0 * ORC_PI / (xmax - xmin);
    for (int j = 0; j < n; ++j) {
        const int shifted = (2 * j < n) ? j : j - n;
        k[j] = dk * shifted;
    }
}

/* ---- radix-2 FFT (core/src/fft.ixx) ------------------------------------- */

/* core/src/fft.ixx:29-37: direct twiddles w[j] = e^{-2 pi i j / n} */
void orc_fft_twiddles(size_t n, double* w) {
    const double ang = -2.0 * ORC_PI / (double)n;
    for (size_t j = 0; j < n / 2; ++j) {
        const double th = ang * (double)j;
        w[2 * j] = cos(th);
        w[2 * j + 1] = sin(th);
    }
}

static int is_pow2(size_t n) { return (n & (n - 1)) == 0; }

/* core/src/fft.ixx:39-73: bit reversal then butterflies, stage len uses w[j*n/len] */
int orc_fft_line(double* a, size_t n, const double* w) {
    if (!is_pow2(n)) return -1;
    if (n < 2) return 0;
    for (size_t i = 1, j = 0; i < n; ++i) {
        size_t bit = n >> 1;
        for (; j & bit; bit >>= 1) j ^= bit;
        j ^= bit;
        if (i < j) {
            double t0 = a[2 * i], t1 = a[2 * i + 1];
            a[2 * i] = a[2 * j];
            a[2 * i + 1] = a[2 * j + 1];
            a[2 * j] = t0;
            a[2 * j + 1] = t1;
        }
    }
    for (size_t len = 2; len <= n; len <<= 1) {
        const size_t stride = n / len;
        for (size_t i = 0; i < n; i += len) {
            for (size_t j = 0; j < len / 2; ++j) {
                const size_t p = i + j, q = i + j + len / 2;
                const double ur = a[2 * p], ui = a[2 * p + 1];
                double vr, vi;
                cmul(a[2 * q], a[2 * q + 1], w[2 * (j * stride)], w[2 * (j * stride) + 1], &vr, &vi);
                a[2 * p] = ur + vr;
                a[2 * p + 1] = ui + vi;
                a[2 * q] = ur - vr;
                a[2 * q + 1] = ui - vi;
            }
        }
    }
    return 0;
}

/* core/src/fft.ixx:75-84: twiddles rebuilt on every call */
int orc_fft1d(double* a, size_t n) {
    if (!is_pow2(n)) return -1;
    if (n < 2) return 0;
    double* w = (double*)malloc(sizeof(double) * n);
    orc_fft_twiddles(n, w);
    int rc = orc_fft_line(a, n, w);
    free(w);
    return rc;
}

/* core/src/fft.ixx:89-98: ifft = conj(fft(conj(x))) * (1/N), 1/N multiplied */
int orc_ifft1d(double* a, size_t n) {
    if (!is_pow2(n)) return -1;
    for (size_t i = 0; i < n; ++i) a[2 * i + 1] = -a[2 * i + 1];
    int rc = orc_fft1d(a, n);
    const double inv = 1.0 / (double)n;
    for (size_t i = 0; i < n; ++i) {
        a[2 * i] = inv * a[2 * i];
        a[2 * i + 1] = inv * -a[2 * i + 1];
    }
    return rc;
}

/* Per-axis line FFT over a 3D x-fastest array; each line is independent, so
 * the threaded loop is bitwise identical to the serial one
 * (core/src/fft.ixx:102-173).  `line_fn` is the 1D transform to apply. */
typedef int (*line_fft_fn)(double* line, size_t n, const double* w, int sign);

static int pow2_line(double* line, size_t n, const double* w, int sign) {
    (void)sign;
    return orc_fft_line(line, n, w);
}

static int axis_pass(double* a, const int d[3], int axis, line_fft_fn fn, const double* w,
                     int sign) {
    const int n = d[axis];
    if (n < 2) return 0;
    const long nx = d[0], ny = d[1], nz = d[2];
    const long stride = axis == 0 ? 1 : (axis == 1 ? nx : nx * ny);
    const long lines = (nx * ny * nz) / n;
    int rc = 0;
#pragma omp parallel
    {
        double* line = (double*)malloc(sizeof(double) * 2 * (size_t)n);
#pragma omp for schedule(static)
        for (long m = 0; m < lines; ++m) {
            long base;
            if (axis == 0) {
                base = m * nx;
            } else if (axis == 1) {
                const long k = m / nx, i = m % nx;
                base = i + nx * ny * k;
            } else {
                base = m; /* m = i + nx*j */
            }
            for (int p = 0; p < n; ++p) {
                line[2 * p] = a[2 * (base + p * stride)];
                line[2 * p + 1] = a[2 * (base + p * stride) + 1];
            }
            if (fn(line, (size_t)n, w, sign) != 0) rc = -1;
            for (int p = 0; p < n; ++p) {
                a[2 * (base + p * stride)] = line[2 * p];
                a[2 * (base + p * stride) + 1] = line[2 * p + 1];
            }
        }
        free(line);
    }
    return rc;
}

/* core/src/fft.ixx:102-173: validate up front, then x, y, z lines */
int orc_fft3d(double* a, int nx, int ny, int nz) {
    if (!is_pow2((size_t)nx) || !is_pow2((size_t)ny) || !is_pow2((size_t)nz)) return -1;
    const int d[3] = {nx, ny, nz};
    for (int axis = 0; axis < 3; ++axis) {
        const size_t n = (size_t)d[axis];
        double* w = (double*)malloc(sizeof(double) * (n + 2));
        orc_fft_twiddles(n, w);
        axis_pass(a, d, axis, pow2_line, w, -1);
        free(w);
    }
    return 0;
}

/* core/src/fft.ixx:177-188: parallel conj, fft, parallel inv*conj */
int orc_ifft3d(double* a, int nx, int ny, int nz) {
    if (!is_pow2((size_t)nx) || !is_pow2((size_t)ny) || !is_pow2((size_t)nz)) return -1;
    const long n = (long)nx * ny * nz;
#pragma omp parallel for schedule(static)
    for (long i = 0; i < n; ++i) a[2 * i + 1] = -a[2 * i + 1];
    orc_fft3d(a, nx, ny, nz);
    const double inv = 1.0 / (double)(int)n;
#pragma omp parallel for schedule(static)
    for (long i = 0; i < n; ++i) {
        a[2 * i] = inv * a[2 * i];
        a[2 * i + 1] = inv * -a[2 * i + 1];
    }
    return 0;
}

/* ---- mixed-radix FFT (no reference counterpart) ------------------------- */

/* O(n^2) DFT with sign convention X_k = sum x_n e^{sign 2 pi i k n / N};
 * the angle uses the exact integer (k*n mod N). */
void orc_dft_brute(const double* in, double* out, size_t n, int sign) {
    for (size_t k = 0; k < n; ++k) {
        double sr = 0.0, si = 0.0;
        for (size_t m = 0; m < n; ++m) {
            const double th = sign * 2.0 * ORC_PI * (double)((k * m) % n) / (double)n;
            double pr, pi;
            cmul(in[2 * m], in[2 * m + 1], cos(th), sin(th), &pr, &pi);
            sr += pr;
            si += pi;
        }
        out[2 * k] = sr;
        out[2 * k + 1] = si;
    }
}

/* Recursive decimation in time by the smallest prime factor p:
 *   X[k] = sum_q W_n^{q k} FFT_{n/p}(x[q::p])[k mod n/p].
 * The factor of level `depth` (size n) is W_n^j with j = (q k) mod n, taken
 * from that level's table tab[depth][j] = (cos th, sin th),
 * th = sign 2 pi j / n: the same expression, so the same bits, as forming
 * cos and sin of it term by term. */
typedef struct {
    int levels;
    size_t n[64];
    double* tab[64];
    int sign;
} any_plan;

static void any_plan_make(any_plan* pl, size_t n, int sign) {
    pl->levels = 0;
    pl->sign = sign;
    while (n > 1) {
        size_t p = 2;
        while (n % p) ++p;
        double* t = (double*)malloc(sizeof(double) * 2 * n);
        for (size_t j = 0; j < n; ++j) {
            const double th = sign * 2.0 * ORC_PI * (double)j / (double)n;
            t[2 * j] = cos(th);
            t[2 * j + 1] = sin(th);
        }
        pl->n[pl->levels] = n;
        pl->tab[pl->levels] = t;
        ++pl->levels;
        n /= p;
    }
}

static void any_plan_free(any_plan* pl) {
    for (int l = 0; l < pl->levels; ++l) free(pl->tab[l]);
    pl->levels = 0;
}

static void fft_rec(const double* in, size_t istride, double* out, size_t n, const any_plan* pl,
                    int depth, double* scratch) {
    if (n == 1) {
        out[0] = in[0];
        out[1] = in[1];
        return;
    }
    size_t p = 2;
    while (n % p) ++p;
    const size_t m = n / p;
    const double* w = pl->tab[depth];
    /* sub-transforms into out[q*m .. q*m+m) */
    for (size_t q = 0; q < p; ++q)
        fft_rec(in + 2 * q * istride, istride * p, out + 2 * q * m, m, pl, depth + 1,
                scratch + 2 * n);
    for (size_t k = 0; k < n; ++k) {
        double sr = 0.0, si = 0.0;
        for (size_t q = 0; q < p; ++q) {
            const size_t j = (q * k) % n;
            double pr, pi;
            const double* s = out + 2 * (q * m + (k % m));
            cmul(s[0], s[1], w[2 * j], w[2 * j + 1], &pr, &pi);
            sr += pr;
            si += pi;
        }
        scratch[2 * k] = sr;
        scratch[2 * k + 1] = si;
    }
    memcpy(out, scratch, sizeof(double) * 2 * n);
}

static void fft_any_planned(double* a, size_t n, const any_plan* pl, double* tmp,
                            double* scratch) {
    memcpy(tmp, a, sizeof(double) * 2 * n);
    fft_rec(tmp, 1, a, n, pl, 0, scratch);
}

int orc_fft_any(double* a, size_t n, int sign) {
    if (n < 2) return 0;
    any_plan pl;
    any_plan_make(&pl, n, sign);
    /* scratch depth: sum over recursion levels of 2*n_level <= 4n */
    double* tmp = (double*)malloc(sizeof(double) * 2 * n);
    double* scratch = (double*)malloc(sizeof(double) * 8 * n + 16);
    fft_any_planned(a, n, &pl, tmp, scratch);
    free(scratch);
    free(tmp);
    any_plan_free(&pl);
    return 0;
}

/* A 3D pass of the mixed-radix transform: one plan per axis (passed as the
 * line function's table pointer), per-thread work buffers. */
static int any_line(double* line, size_t n, const double* w, int sign) {
    (void)sign;
    const any_plan* pl = (const any_plan*)(const void*)w;
    double* tmp = (double*)malloc(sizeof(double) * 2 * n);
    double* scratch = (double*)malloc(sizeof(double) * 8 * n + 16);
    fft_any_planned(line, n, pl, tmp, scratch);
    free(scratch);
    free(tmp);
    return 0;
}

int orc_fft3d_any(double* a, int nx, int ny, int nz, int sign) {
    const int d[3] = {nx, ny, nz};
    for (int axis = 0; axis < 3; ++axis) {
        if (d[axis] < 2) continue;
        any_plan pl;
        any_plan_make(&pl, (size_t)d[axis], sign);
        axis_pass(a, d, axis, any_line, (const double*)(const void*)&pl, sign);
        any_plan_free(&pl);
    }
    return 0;
}

/* ---- split-operator propagator (core/src/propagator.ixx) ---------------- */

/* core/src/propagator.ixx:23-40 */
void orc_so1d_tables(const double* v, int n, double xmin, double xmax, double dt,
                     double* half_v, double* kinetic) {
    for (int i = 0; i < n; ++i) {
        const double th = -0.5 * v[i] * dt;
        half_v[2 * i] = cos(th);
        half_v[2 * i + 1] = sin(th);
    }
    double* k = (double*)malloc(sizeof(double) * (size_t)n);
    orc_wavenumbers(n, xmin, xmax, k);
    for (int j = 0; j < n; ++j) {
        const double th = -0.5 * k[j] * k[j] * dt;
        kinetic[2 * j] = cos(th);
        kinetic[2 * j + 1] = sin(th);
    }
    free(k);
}

/* core/src/propagator.ixx:56-61 and 123-129: a[i] = a[i] * phase[i] */
static void apply_phase(const double* phase, double* a, long n) {
#pragma omp parallel for schedule(static)
    for (long i = 0; i < n; ++i) {
        double r, im;
        cmul(a[2 * i], a[2 * i + 1], phase[2 * i], phase[2 * i + 1], &r, &im);
        a[2 * i] = r;
        a[2 * i + 1] = im;
    }
}

/* core/src/propagator.ixx:44-53 */
int orc_so1d_step(double* psi, int n, const double* half_v, const double* kinetic,
                  int nsteps) {
    if (!is_pow2((size_t)n)) return -1;
    for (int s = 0; s < nsteps; ++s) {
        apply_phase(half_v, psi, n);
        orc_fft1d(psi, (size_t)n);
        apply_phase(kinetic, psi, n);
        orc_ifft1d(psi, (size_t)n);
        apply_phase(half_v, psi, n);
    }
    return 0;
}

/* core/src/propagator.ixx:72-101: full N-entry kinetic table */
void orc_so3d_tables(const double* v, const int d[3], const double lo[3], const double hi[3],
                     double dt, double mass, double* half_v, double* kinetic) {
    const long n = (long)d[0] * d[1] * d[2];
    for (long i = 0; i < n; ++i) {
        const double th = -0.5 * v[i] * dt;
        half_v[2 * i] = cos(th);
        half_v[2 * i + 1] = sin(th);
    }
    double* kx = (double*)malloc(sizeof(double) * (size_t)d[0]);
    double* ky = (double*)malloc(sizeof(double) * (size_t)d[1]);
    double* kz = (double*)malloc(sizeof(double) * (size_t)d[2]);
    orc_wavenumbers(d[0], lo[0], hi[0], kx);
    orc_wavenumbers(d[1], lo[1], hi[1], ky);
    orc_wavenumbers(d[2], lo[2], hi[2], kz);
#pragma omp parallel for schedule(static)
    for (int k = 0; k < d[2]; ++k) {
        for (int j = 0; j < d[1]; ++j) {
            for (int i = 0; i < d[0]; ++i) {
                const double kxx = kx[i], kyy = ky[j], kzz = kz[k];
                const double th = -0.5 * (kxx * kxx + kyy * kyy + kzz * kzz) / mass * dt;
                const long f = i + (long)d[0] * (j + (long)d[1] * k);
                kinetic[2 * f] = cos(th);
                kinetic[2 * f + 1] = sin(th);
            }
        }
    }
    free(kx);
    free(ky);
    free(kz);
}

/* core/src/propagator.ixx:110-119 */
int orc_so3d_step(double* psi, const int d[3], const double* half_v, const double* kinetic,
                  int nsteps) {
    if (!is_pow2((size_t)d[0]) || !is_pow2((size_t)d[1]) || !is_pow2((size_t)d[2])) return -1;
    const long n = (long)d[0] * d[1] * d[2];
    for (int s = 0; s < nsteps; ++s) {
        apply_phase(half_v, psi, n);
        orc_fft3d(psi, d[0], d[1], d[2]);
        apply_phase(kinetic, psi, n);
        orc_ifft3d(psi, d[0], d[1], d[2]);
        apply_phase(half_v, psi, n);
    }
    return 0;
}

/* Same Strang step with the mixed-radix FFT; the inverse is the e^{+} transform
 * times 1/N (equal in exact arithmetic to the conjugation identity). */
int orc_so3d_step_any(double* psi, const int d[3], const double* half_v,
                      const double* kinetic, int nsteps) {
    const long n = (long)d[0] * d[1] * d[2];
    const double inv = 1.0 / (double)n;
    for (int s = 0; s < nsteps; ++s) {
        apply_phase(half_v, psi, n);
        orc_fft3d_any(psi, d[0], d[1], d[2], -1);
        apply_phase(kinetic, psi, n);
        orc_fft3d_any(psi, d[0], d[1], d[2], +1);
#pragma omp parallel for schedule(static)
        for (long i = 0; i < n; ++i) {
            psi[2 * i] = inv * psi[2 * i];
            psi[2 * i + 1] = inv * psi[2 * i + 1];
        }
        apply_phase(half_v, psi, n);
    }
    return 0;
}

/* scenario/src/line1d_director.ixx:205-209: psi[i] *= mask[i] (real mask) */
void orc_apply_mask(double* psi, const double* mask, size_t n) {
#pragma omp parallel for schedule(static)
    for (long i = 0; i < (long)n; ++i) {
        psi[2 * i] = psi[2 * i] * mask[i];
        psi[2 * i + 1] = psi[2 * i + 1] * mask[i];
    }
}

/* ---- field reductions (core/src/field.ixx) ------------------------------ */

/* core/src/field.ixx:33-39 and 70-76: serial sum of |z|^2, times the cell volume */
double orc_norm_sq(const double* psi, size_t n, double cell_volume) {
    double acc = 0.0;
    for (size_t i = 0; i < n; ++i) acc += psi[2 * i] * psi[2 * i] + psi[2 * i + 1] * psi[2 * i + 1];
    return acc * cell_volume;
}

/* core/src/field.ixx:78-87 (the 1D variant, field.ixx:41-46, has no zero guard;
 * for a non-zero field both are the same operation) */
void orc_normalize(double* psi, size_t n, double cell_volume) {
    const double n2 = orc_norm_sq(psi, n, cell_volume);
    if (n2 <= 0.0) return;
    const double inv = 1.0 / sqrt(n2);
    for (size_t i = 0; i < n; ++i) {
        psi[2 * i] = inv * psi[2 * i];
        psi[2 * i + 1] = inv * psi[2 * i + 1];
    }
}

/* ---- dipole drive (core/src/drive.ixx) ---------------------------------- */

/* core/src/drive.ixx:25-40: psi *= exp(-i theta (axis . r)),
 * theta = amplitude cos(omega t) dt / 2 */
void orc_dipole_halfkick(double* psi, const int d[3], const double lo[3], const double hi[3],
                         const double axis[3], double amplitude, double omega, double t,
                         double dt) {
    const double theta = amplitude * cos(omega * t) * 0.5 * dt;
    for (int k = 0; k < d[2]; ++k)
        for (int j = 0; j < d[1]; ++j)
            for (int i = 0; i < d[0]; ++i) {
                const double u = axis[0] * coord(lo[0], hi[0], d[0], i) +
                                 axis[1] * coord(lo[1], hi[1], d[1], j) +
                                 axis[2] * coord(lo[2], hi[2], d[2], k);
                const double ang = -theta * u;
                const long f = i + (long)d[0] * (j + (long)d[1] * k);
                double r, im;
                cmul(psi[2 * f], psi[2 * f + 1], cos(ang), sin(ang), &r, &im);
                psi[2 * f] = r;
                psi[2 * f + 1] = im;
            }
}

/* core/src/drive.ixx:52-66: each step is halfkick(t), the SplitOperator3D
 * step with its tables, halfkick(t + dt), t = t0 + s dt.  any_size: the
 * mixed-radix FFT (no reference counterpart). */
int orc_driven_step(double* psi, const int d[3], const double lo[3], const double hi[3],
                    const double* half_v, const double* kinetic, const double axis[3],
                    double amplitude, double omega, double t0, double dt, int nsteps,
                    int any_size) {
    for (int s = 0; s < nsteps; ++s) {
        const double t = t0 + s * dt;
        orc_dipole_halfkick(psi, d, lo, hi, axis, amplitude, omega, t, dt);
        const int rc = any_size ? orc_so3d_step_any(psi, d, half_v, kinetic, 1)
                                : orc_so3d_step(psi, d, half_v, kinetic, 1);
        if (rc != 0) return rc;
        orc_dipole_halfkick(psi, d, lo, hi, axis, amplitude, omega, t + dt, dt);
    }
    return 0;
}

/* ---- uniform magnetic field (core/src/magnetic.ixx, core/src/rotation.ixx) */

/* core/src/rotation.ixx:25-66: Fourier-shift every line along freq_axis by
 * coeff * coord(coord_axis of the line) */
static int axis_shear(double* a, const int d[3], const double lo[3], const double hi[3],
                      int freq_axis, int coord_axis, double coeff) {
    const int s[3] = {1, d[0], d[0] * d[1]};
    const int nf = d[freq_axis], sf = s[freq_axis];
    if (!is_pow2((size_t)nf)) return -1;
    int perp[2], pc = 0;
    for (int q = 0; q < 3; ++q)
        if (q != freq_axis) perp[pc++] = q;
    double* kf = (double*)malloc(sizeof(double) * (size_t)nf);
    orc_wavenumbers(nf, lo[freq_axis], hi[freq_axis], kf);
    double* line = (double*)malloc(sizeof(double) * 2 * (size_t)nf);
    for (int v = 0; v < d[perp[1]]; ++v) {
        for (int u = 0; u < d[perp[0]]; ++u) {
            const int base = u * s[perp[0]] + v * s[perp[1]];
            const int cidx = (coord_axis == perp[0]) ? u : v;
            const double dd = coeff * coord(lo[coord_axis], hi[coord_axis], d[coord_axis], cidx);
            for (int t = 0; t < nf; ++t) {
                line[2 * t] = a[2 * (size_t)(base + t * sf)];
                line[2 * t + 1] = a[2 * (size_t)(base + t * sf) + 1];
            }
            orc_fft1d(line, (size_t)nf);
            for (int t = 0; t < nf; ++t) {
                const double ph = -kf[t] * dd;
                double r, im;
                cmul(line[2 * t], line[2 * t + 1], cos(ph), sin(ph), &r, &im);
                line[2 * t] = r;
                line[2 * t + 1] = im;
            }
            orc_ifft1d(line, (size_t)nf);
            for (int t = 0; t < nf; ++t) {
                a[2 * (size_t)(base + t * sf)] = line[2 * t];
                a[2 * (size_t)(base + t * sf) + 1] = line[2 * t + 1];
            }
        }
    }
    free(line);
    free(kf);
    return 0;
}

/* core/src/rotation.ixx:73-84: Paeth three-shear rotation about `axis` */
int orc_rotate_axis(double* psi, const int d[3], const double lo[3], const double hi[3],
                    int axis, double theta) {
    const int b = (axis + 1) % 3, c = (axis + 2) % 3;
    const double t = tan(0.5 * theta);
    const double s = sin(theta);
    if (axis_shear(psi, d, lo, hi, b, c, -t)) return -1;
    if (axis_shear(psi, d, lo, hi, c, b, s)) return -1;
    return axis_shear(psi, d, lo, hi, b, c, -t);
}

/* core/src/magnetic.ixx:50-73: v + B^2/8 |r_perp|^2 */
void orc_magnetic_veff(const double* v, const int d[3], const double lo[3], const double hi[3],
                       double bfield, int axis, double* out) {
    const double c = bfield * bfield / 8.0;
    for (int k = 0; k < d[2]; ++k)
        for (int j = 0; j < d[1]; ++j)
            for (int i = 0; i < d[0]; ++i) {
                const long f = i + (long)d[0] * (j + (long)d[1] * k);
                out[f] = v[f];
                if (c != 0.0) {
                    const double co[3] = {coord(lo[0], hi[0], d[0], i), coord(lo[1], hi[1], d[1], j),
                                          coord(lo[2], hi[2], d[2], k)};
                    double perp2 = 0.0;
                    for (int a = 0; a < 3; ++a)
                        if (a != axis) perp2 += co[a] * co[a];
                    out[f] += c * perp2;
                }
            }
}

/* core/src/magnetic.ixx:34-46: R(a), SplitOperator3D step (tables of the
 * effective potential), R(a); a = (B/2)(dt/2) */
int orc_magnetic_step(double* psi, const int d[3], const double lo[3], const double hi[3],
                      const double* half_v, const double* kinetic, double dt, double bfield,
                      int axis, int nsteps) {
    const double half_angle = 0.5 * bfield * (0.5 * dt);
    for (int s = 0; s < nsteps; ++s) {
        if (half_angle != 0.0 && orc_rotate_axis(psi, d, lo, hi, axis, half_angle)) return -1;
        if (orc_so3d_step(psi, d, half_v, kinetic, 1)) return -1;
        if (half_angle != 0.0 && orc_rotate_axis(psi, d, lo, hi, axis, half_angle)) return -1;
    }
    return 0;
}

/* ---- Peierls lattice (core/src/lattice2d.ixx) ------------------------- */

/* Complex helpers with the reference's -fcx-limited-range products. */
static inline void cm(const double* a, const double* b, double* out) {
    const double r = a[0] * b[0] - a[1] * b[1];
    const double i = a[0] * b[1] + a[1] * b[0];
    out[0] = r;
    out[1] = i;
}

/* core/src/lattice2d.ixx:33-58: hopping t = 1/(2 h^2) per axis, the bond
 * mixing constants (x: dt/2, y-even: dt/2, y-odd: dt) and the onsite tables
 * (FD diagonal 2 tx + 2 ty added to V).  consts = {tx, ty, cx, sx, cy, sy,
 * cy2, sy2, rcx, rsx, rcy, rsy, rcy2, rsy2}. */
void orc_lat2d_tables(const double* v, int nx, int ny, double hx, double hy, double dt,
                      double* consts, double* half_v, double* relax_half_v) {
    const double tx = 0.5 / (hx * hx), ty = 0.5 / (hy * hy);
    consts[0] = tx;
    consts[1] = ty;
    consts[2] = cos(tx * 0.5 * dt);
    consts[3] = sin(tx * 0.5 * dt);
    consts[4] = cos(ty * 0.5 * dt);
    consts[5] = sin(ty * 0.5 * dt);
    consts[6] = cos(ty * dt);
    consts[7] = sin(ty * dt);
    consts[8] = cosh(tx * 0.5 * dt);
    consts[9] = sinh(tx * 0.5 * dt);
    consts[10] = cosh(ty * 0.5 * dt);
    consts[11] = sinh(ty * 0.5 * dt);
    consts[12] = cosh(ty * dt);
    consts[13] = sinh(ty * dt);
    for (long i = 0; i < (long)nx * ny; ++i) {
        const double e0 = v[i] + 2.0 * tx + 2.0 * ty;
        const double th = -0.5 * e0 * dt;
        half_v[2 * i] = cos(th);
        half_v[2 * i + 1] = sin(th);
        relax_half_v[i] = exp(-0.5 * e0 * dt);
    }
}

/* core/src/lattice2d.ixx:63-91: string-gauge solenoid on the x-links;
 * link_x is reset to 1 first.  core/src/lattice2d.ixx:96-105: uniform field,
 * Landau gauge A_x = B y. */
void orc_lat2d_solenoid(double* link_x, int nx, int ny, double xmin, double xmax, double ymin,
                        double ymax, double phi, double xs, double ys, int cut_up) {
    for (long i = 0; i < (long)nx * ny; ++i) {
        link_x[2 * i] = 1.0;
        link_x[2 * i + 1] = 0.0;
    }
    int is = -1, js = -1;
    for (int i = 0; i + 1 < nx; ++i)
        if (coord(xmin, xmax, nx, i) <= xs && xs < coord(xmin, xmax, nx, i + 1)) is = i;
    for (int j = 0; j + 1 < ny; ++j)
        if (coord(ymin, ymax, ny, j) <= ys && ys < coord(ymin, ymax, ny, j + 1)) js = j;
    if (is < 0 || js < 0) return;
    const double up[2] = {cos(phi), -sin(phi)};
    const double dn[2] = {cos(phi), sin(phi)};
    if (cut_up) {
        for (int j = js + 1; j < ny; ++j) {
            link_x[2 * ((long)j * nx + is)] = up[0];
            link_x[2 * ((long)j * nx + is) + 1] = up[1];
        }
    } else {
        for (int j = 0; j <= js; ++j) {
            link_x[2 * ((long)j * nx + is)] = dn[0];
            link_x[2 * ((long)j * nx + is) + 1] = dn[1];
        }
    }
}

void orc_lat2d_uniform_field(double* link_x, int nx, int ny, double xmin, double xmax,
                             double ymin, double ymax, double b) {
    const double bh = b * orc_spacing(xmin, xmax, nx);
    for (int j = 0; j < ny; ++j) {
        const double th = -bh * coord(ymin, ymax, ny, j);
        const double c = cos(th), s = sin(th);
        for (int i = 0; i < nx; ++i) {
            link_x[2 * ((long)j * nx + i)] = c;
            link_x[2 * ((long)j * nx + i) + 1] = s;
        }
    }
}

/* core/src/lattice2d.ixx:178-197 (x) and 199-216 (y): one bond parity group,
 * a_i = c a_i + mix u a_k, a_k = mix conj(u) a_i + c a_k */
static void lat_bond(double* a, long i, long k, const double* u, double c, const double* mix) {
    const double pa[2] = {a[2 * i], a[2 * i + 1]};
    const double pb[2] = {a[2 * k], a[2 * k + 1]};
    double mu[2], t[2], mcu[2], t2[2];
    cm(mix, u, mu);
    cm(mu, pb, t);
    a[2 * i] = c * pa[0] + t[0];
    a[2 * i + 1] = c * pa[1] + t[1];
    const double uc[2] = {u[0], -u[1]};
    cm(mix, uc, mcu);
    cm(mcu, pa, t2);
    a[2 * k] = t2[0] + c * pb[0];
    a[2 * k + 1] = t2[1] + c * pb[1];
}

static void lat_sweep_x(double* a, int nx, int ny, const double* link_x, int parity, double c,
                        const double* mix) {
    for (int j = 0; j < ny; ++j) {
        const long row = (long)j * nx;
        for (int i = parity; i + 1 < nx; i += 2)
            lat_bond(a, row + i, row + i + 1, link_x + 2 * (row + i), c, mix);
    }
}

static void lat_sweep_y(double* a, int nx, int ny, const double* link_y, int parity, double c,
                        const double* mix) {
    for (int j = parity; j + 1 < ny; j += 2) {
        const long lo = (long)j * nx, hi = (long)(j + 1) * nx;
        for (int i = 0; i < nx; ++i) lat_bond(a, lo + i, hi + i, link_y + 2 * (lo + i), c, mix);
    }
}

static void lat_phase(double* a, long n, const double* table) {
    for (long i = 0; i < n; ++i) {
        double t[2];
        cm(a + 2 * i, table + 2 * i, t);
        a[2 * i] = t[0];
        a[2 * i + 1] = t[1];
    }
}

/* core/src/lattice2d.ixx:218-232 (step) and 234-249 (relax: cosh/sinh
 * mixing, real onsite decay as a complex (w, 0) factor, then normalise
 * with cell volume dv).  relax_half_v is real; link_y is all ones. */
int orc_lat2d_evolve(double* psi, int nx, int ny, const double* consts, const double* half_v,
                     const double* relax_half_v, const double* link_x, int relax, double dv,
                     int nsteps) {
    const long n = (long)nx * ny;
    double* link_y = (double*)malloc(sizeof(double) * 2 * (size_t)n);
    for (long i = 0; i < n; ++i) {
        link_y[2 * i] = 1.0;
        link_y[2 * i + 1] = 0.0;
    }
    double* rtab = NULL;
    if (relax) {
        rtab = (double*)malloc(sizeof(double) * 2 * (size_t)n);
        for (long i = 0; i < n; ++i) {
            rtab[2 * i] = relax_half_v[i];
            rtab[2 * i + 1] = 0.0;
        }
    }
    const double* tab = relax ? rtab : half_v;
    const double cx = consts[relax ? 8 : 2], cy = consts[relax ? 10 : 4],
                 cy2 = consts[relax ? 12 : 6];
    double mx[2], my[2], my2[2];
    if (relax) {
        mx[0] = consts[9]; mx[1] = 0.0;
        my[0] = consts[11]; my[1] = 0.0;
        my2[0] = consts[13]; my2[1] = 0.0;
    } else {
        mx[0] = 0.0; mx[1] = consts[3];
        my[0] = 0.0; my[1] = consts[5];
        my2[0] = 0.0; my2[1] = consts[7];
    }
    for (int s = 0; s < nsteps; ++s) {
        lat_phase(psi, n, tab);
        lat_sweep_x(psi, nx, ny, link_x, 0, cx, mx);
        lat_sweep_x(psi, nx, ny, link_x, 1, cx, mx);
        lat_sweep_y(psi, nx, ny, link_y, 0, cy, my);
        lat_sweep_y(psi, nx, ny, link_y, 1, cy2, my2);
        lat_sweep_y(psi, nx, ny, link_y, 0, cy, my);
        lat_sweep_x(psi, nx, ny, link_x, 1, cx, mx);
        lat_sweep_x(psi, nx, ny, link_x, 0, cx, mx);
        lat_phase(psi, n, tab);
        if (relax) orc_normalize(psi, (size_t)n, dv);
    }
    free(link_y);
    free(rtab);
    return 0;
}

/* core/src/lattice2d.ixx:251-274: <H> = hops + onsite, normalised */
double orc_lat2d_energy(const double* psi, int nx, int ny, const double* v, const double* consts,
                        const double* link_x) {
    const double tx = consts[0], ty = consts[1];
    double e = 0.0, den = 0.0;
    for (int j = 0; j < ny; ++j) {
        for (int i = 0; i < nx; ++i) {
            const long f = (long)j * nx + i;
            const double z[2] = {psi[2 * f], psi[2 * f + 1]};
            const double w = z[0] * z[0] + z[1] * z[1];
            den += w;
            e += (v[f] + 2.0 * tx + 2.0 * ty) * w;
            const double zc[2] = {z[0], -z[1]};
            if (i + 1 < nx) {
                double t[2], t2[2];
                cm(zc, link_x + 2 * f, t);
                cm(t, psi + 2 * (f + 1), t2);
                e += -tx * 2.0 * t2[0];
            }
            if (j + 1 < ny) {
                const double one[2] = {1.0, 0.0};
                double t[2], t2[2];
                cm(zc, one, t);
                cm(t, psi + 2 * (f + nx), t2);
                e += -ty * 2.0 * t2[0];
            }
        }
    }
    return e / den;
}

/* ---- imaginary-time relaxation (core/src/imaginary_time.ixx) ---------- */

/* core/src/imaginary_time.ixx:24-40 */
void orc_it1d_tables(const double* v, int n, double xmin, double xmax, double dtau,
                     double* half_w, double* kin_w) {
    for (int i = 0; i < n; ++i) half_w[i] = exp(-0.5 * v[i] * dtau);
    double* k = (double*)malloc(sizeof(double) * (size_t)n);
    orc_wavenumbers(n, xmin, xmax, k);
    for (int j = 0; j < n; ++j) kin_w[j] = exp(-0.5 * k[j] * k[j] * dtau);
    free(k);
}

/* core/src/imaginary_time.ixx:67-94: full N-entry kinetic weight */
void orc_it3d_tables(const double* v, const int d[3], const double lo[3], const double hi[3],
                     double dtau, double mass, double* half_w, double* kin_w) {
    const long n = (long)d[0] * d[1] * d[2];
    for (long i = 0; i < n; ++i) half_w[i] = exp(-0.5 * v[i] * dtau);
    double* kx = (double*)malloc(sizeof(double) * (size_t)d[0]);
    double* ky = (double*)malloc(sizeof(double) * (size_t)d[1]);
    double* kz = (double*)malloc(sizeof(double) * (size_t)d[2]);
    orc_wavenumbers(d[0], lo[0], hi[0], kx);
    orc_wavenumbers(d[1], lo[1], hi[1], ky);
    orc_wavenumbers(d[2], lo[2], hi[2], kz);
#pragma omp parallel for schedule(static)
    for (int k = 0; k < d[2]; ++k)
        for (int j = 0; j < d[1]; ++j)
            for (int i = 0; i < d[0]; ++i) {
                const double kxx = kx[i], kyy = ky[j], kzz = kz[k];
                kin_w[i + (long)d[0] * (j + (long)d[1] * k)] =
                    exp(-0.5 * (kxx * kxx + kyy * kyy + kzz * kzz) / mass * dtau);
            }
    free(kx);
    free(ky);
    free(kz);
}

/* core/src/imaginary_time.ixx:57-62 and 136-141: a[i] = weight[i] * a[i] */
static void apply_weight(const double* w, double* a, long n) {
#pragma omp parallel for schedule(static)
    for (long i = 0; i < n; ++i) {
        a[2 * i] = w[i] * a[2 * i];
        a[2 * i + 1] = w[i] * a[2 * i + 1];
    }
}

/* core/src/imaginary_time.ixx:43-53 (1D normalize: field.ixx:41-46).
 * last_norm (nullable) receives the norm of the last step before its
 * renormalisation. */
int orc_it1d_relax(double* psi, int n, double dx, const double* half_w, const double* kin_w,
                   int nsteps, double* last_norm) {
    if (!is_pow2((size_t)n)) return -1;
    for (int s = 0; s < nsteps; ++s) {
        apply_weight(half_w, psi, n);
        orc_fft1d(psi, (size_t)n);
        apply_weight(kin_w, psi, n);
        orc_ifft1d(psi, (size_t)n);
        apply_weight(half_w, psi, n);
        const double n2 = orc_norm_sq(psi, (size_t)n, dx);
        if (last_norm) *last_norm = n2;
        const double inv = 1.0 / sqrt(n2);
        for (int i = 0; i < 2 * n; ++i) psi[i] = inv * psi[i];
    }
    return 0;
}

/* core/src/field.ixx:88-99: <a|b> = sum conj(a) b dV, real and imaginary
 * parts summed separately in index order */
void orc_inner_product(const double* a, const double* b, size_t n, double dv, double* out) {
    double re = 0.0, im = 0.0;
    for (size_t i = 0; i < n; ++i) {
        /* conj(a) * b with -fcx-limited-range: plain four-product form */
        const double ar = a[2 * i], ai = -a[2 * i + 1], br = b[2 * i], bi = b[2 * i + 1];
        re += ar * br - ai * bi;
        im += ar * bi + ai * br;
    }
    out[0] = re * dv;
    out[1] = im * dv;
}

/* core/src/imaginary_time.ixx:112-122 (relax) and 124-141 (relax_deflated;
 * nlower = 0 is relax).  any_size: the mixed-radix FFT (no reference
 * counterpart) for non-power-of-two axes. */
int orc_it3d_relax(double* psi, const int d[3], double dv, const double* half_w,
                   const double* kin_w, const double* const* lower, int nlower, int nsteps,
                   int any_size, double* last_norm) {
    const long n = (long)d[0] * d[1] * d[2];
    if (!any_size &&
        (!is_pow2((size_t)d[0]) || !is_pow2((size_t)d[1]) || !is_pow2((size_t)d[2])))
        return -1;
    const double inv_n = 1.0 / (double)n;
    for (int s = 0; s < nsteps; ++s) {
        apply_weight(half_w, psi, n);
        if (any_size) orc_fft3d_any(psi, d[0], d[1], d[2], -1);
        else orc_fft3d(psi, d[0], d[1], d[2]);
        apply_weight(kin_w, psi, n);
        if (any_size) {
            orc_fft3d_any(psi, d[0], d[1], d[2], +1);
            for (long i = 0; i < 2 * n; ++i) psi[i] = inv_n * psi[i];
        } else {
            orc_ifft3d(psi, d[0], d[1], d[2]);
        }
        apply_weight(half_w, psi, n);
        for (int j = 0; j < nlower; ++j) {
            double c[2];
            orc_inner_product(lower[j], psi, (size_t)n, dv, c);
            for (long i = 0; i < n; ++i) {
                /* p[i] - c * q[i] */
                const double qr = lower[j][2 * i], qi = lower[j][2 * i + 1];
                psi[2 * i] = psi[2 * i] - (c[0] * qr - c[1] * qi);
                psi[2 * i + 1] = psi[2 * i + 1] - (c[0] * qi + c[1] * qr);
            }
        }
        const double n2 = orc_norm_sq(psi, (size_t)n, dv);
        if (last_norm) *last_norm = n2;
        orc_normalize(psi, (size_t)n, dv);
    }
    return 0;
}

/* core/src/potential.ixx:25-33 */
void orc_soft_coulomb_1d(double xmin, double xmax, int n, double z, double a, double x0,
                         double* v) {
    for (int i = 0; i < n; ++i) {
        const double dx = coord(xmin, xmax, n, i) - x0;
        v[i] = -z / sqrt(dx * dx + a * a);
    }
}

/* core/src/potential.ixx:51-65 */
void orc_soft_coulomb_3d(const int d[3], const double lo[3], const double hi[3], double z,
                         double a, const double c[3], double* v) {
    for (int k = 0; k < d[2]; ++k)
        for (int j = 0; j < d[1]; ++j)
            for (int i = 0; i < d[0]; ++i) {
                const double dx = coord(lo[0], hi[0], d[0], i) - c[0];
                const double dy = coord(lo[1], hi[1], d[1], j) - c[1];
                const double dz = coord(lo[2], hi[2], d[2], k) - c[2];
                v[i + (long)d[0] * (j + (long)d[1] * k)] =
                    -z / sqrt(dx * dx + dy * dy + dz * dz + a * a);
            }
}

/* ---- builders (core/src/potential.ixx, core/src/wavepacket.ixx) --------- */

static inline double cell_volume3(const double lo[3], const double hi[3], const int d[3]) {
    return orc_spacing(lo[0], hi[0], d[0]) * orc_spacing(lo[1], hi[1], d[1]) *
           orc_spacing(lo[2], hi[2], d[2]);
}

/* core/src/potential.ixx:16-23 */
void orc_harmonic_1d(double xmin, double xmax, int n, double omega, double x0, double* v) {
    for (int i = 0; i < n; ++i) {
        const double dx = coord(xmin, xmax, n, i) - x0;
        v[i] = 0.5 * omega * omega * dx * dx;
    }
}

/* core/src/potential.ixx:34-48 */
void orc_harmonic_3d(const int d[3], const double lo[3], const double hi[3], double omega,
                     const double c[3], double* v) {
    for (int k = 0; k < d[2]; ++k)
        for (int j = 0; j < d[1]; ++j)
            for (int i = 0; i < d[0]; ++i) {
                const double dx = coord(lo[0], hi[0], d[0], i) - c[0];
                const double dy = coord(lo[1], hi[1], d[1], j) - c[1];
                const double dz = coord(lo[2], hi[2], d[2], k) - c[2];
                v[i + (long)d[0] * (j + (long)d[1] * k)] =
                    0.5 * omega * omega * (dx * dx + dy * dy + dz * dz);
            }
}

/* core/src/potential.ixx:102-112: half-open [x_lo, x_hi) */
void orc_barrier_1d(double xmin, double xmax, int n, double v0, double x_lo, double x_hi,
                    double* v) {
    for (int i = 0; i < n; ++i) {
        const double x = coord(xmin, xmax, n, i);
        v[i] = (x >= x_lo && x < x_hi) ? v0 : 0.0;
    }
}

/* core/src/potential.ixx:114-132 */
void orc_barrier_3d(const int d[3], const double lo[3], const double hi[3], double v0,
                    double x_lo, double x_hi, double* v) {
    const long n = (long)d[0] * d[1] * d[2];
    for (long f = 0; f < n; ++f) v[f] = 0.0;
    for (int i = 0; i < d[0]; ++i) {
        const double x = coord(lo[0], hi[0], d[0], i);
        if (x < x_lo || x >= x_hi) continue;
        for (int k = 0; k < d[2]; ++k)
            for (int j = 0; j < d[1]; ++j) v[i + (long)d[0] * (j + (long)d[1] * k)] = v0;
    }
}

/* core/src/potential.ixx:134-154: bare -Z/r, nucleus cell -Z*2.3800774/h */
void orc_regularized_coulomb_3d(const int d[3], const double lo[3], const double hi[3],
                                double z, const double c[3], double* v) {
    const double h = orc_spacing(lo[0], hi[0], d[0]);
    const double center = -z * 2.3800774 / h;
    for (int k = 0; k < d[2]; ++k)
        for (int j = 0; j < d[1]; ++j)
            for (int i = 0; i < d[0]; ++i) {
                const double dx = coord(lo[0], hi[0], d[0], i) - c[0];
                const double dy = coord(lo[1], hi[1], d[1], j) - c[1];
                const double dz = coord(lo[2], hi[2], d[2], k) - c[2];
                const double r = sqrt(dx * dx + dy * dy + dz * dz);
                v[i + (long)d[0] * (j + (long)d[1] * k)] = (r < 1e-6 * h) ? center : -z / r;
            }
}

/* No reference counterpart: the hydrogen workload's V = -Z / max(r, h/2).
 * Pinned by its closed form at the nucleus cell (-2Z/h) and away from it (-Z/r). */
void orc_coulomb_floor_3d(const int d[3], const double lo[3], const double hi[3], double z,
                          const double c[3], double* v) {
    const double h = orc_spacing(lo[0], hi[0], d[0]);
    const double rmin = 0.5 * h;
    for (int k = 0; k < d[2]; ++k)
        for (int j = 0; j < d[1]; ++j)
            for (int i = 0; i < d[0]; ++i) {
                const double dx = coord(lo[0], hi[0], d[0], i) - c[0];
                const double dy = coord(lo[1], hi[1], d[1], j) - c[1];
                const double dz = coord(lo[2], hi[2], d[2], k) - c[2];
                const double r = sqrt(dx * dx + dy * dy + dz * dz);
                v[i + (long)d[0] * (j + (long)d[1] * k)] = -z / (r > rmin ? r : rmin);
            }
}

/* core/src/potential.ixx:194-215 */
void orc_absorbing_mask_1d(double xmin, double xmax, int n, double width, double* m) {
    if (n == 1) {
        m[0] = 1.0;
        return;
    }
    for (int i = 0; i < n; ++i) {
        const double x = coord(xmin, xmax, n, i);
        const double d_lo = x - xmin;
        const double d_hi = xmax - x;
        const double d = d_lo < d_hi ? d_lo : d_hi;
        if (d >= width) {
            m[i] = 1.0;
            continue;
        }
        const double t = d / width;
        const double s = sin(0.5 * 3.14159265358979323846 * t);
        m[i] = s * s;
    }
}

/* core/src/potential.ixx:217-233: separable product mx*my*mz */
void orc_absorbing_mask_3d(const int d[3], const double lo[3], const double hi[3],
                           double width, double* m) {
    double* mx = (double*)malloc(sizeof(double) * (size_t)d[0]);
    double* my = (double*)malloc(sizeof(double) * (size_t)d[1]);
    double* mz = (double*)malloc(sizeof(double) * (size_t)d[2]);
    orc_absorbing_mask_1d(lo[0], hi[0], d[0], width, mx);
    orc_absorbing_mask_1d(lo[1], hi[1], d[1], width, my);
    orc_absorbing_mask_1d(lo[2], hi[2], d[2], width, mz);
    for (int k = 0; k < d[2]; ++k)
        for (int j = 0; j < d[1]; ++j)
            for (int i = 0; i < d[0]; ++i)
                m[i + (long)d[0] * (j + (long)d[1] * k)] = mx[i] * my[j] * mz[k];
    free(mx);
    free(my);
    free(mz);
}

/* core/src/wavepacket.ixx:16-25 */
void orc_gaussian_1d(double xmin, double xmax, int n, double x0, double sigma, double k0,
                     double* psi) {
    const double amp = pow(2.0 * ORC_PI * sigma * sigma, -0.25);
    for (int i = 0; i < n; ++i) {
        const double x = coord(xmin, xmax, n, i);
        const double envelope = amp * exp(-(x - x0) * (x - x0) / (4.0 * sigma * sigma));
        psi[2 * i] = envelope * cos(k0 * x);
        psi[2 * i + 1] = envelope * sin(k0 * x);
    }
    /* core/src/field.ixx:41-46 */
    const double inv = 1.0 / sqrt(orc_norm_sq(psi, (size_t)n, orc_spacing(xmin, xmax, n)));
    for (int i = 0; i < n; ++i) {
        psi[2 * i] = inv * psi[2 * i];
        psi[2 * i + 1] = inv * psi[2 * i + 1];
    }
}

static inline double envelope(double u, double u0, double s) {
    return pow(2.0 * ORC_PI * s * s, -0.25) * exp(-(u - u0) * (u - u0) / (4.0 * s * s));
}

/* core/src/wavepacket.ixx:27-46 */
void orc_gaussian_3d(const int d[3], const double lo[3], const double hi[3], const double r0[3],
                     const double sigma[3], const double k0[3], double* psi) {
    for (int k = 0; k < d[2]; ++k)
        for (int j = 0; j < d[1]; ++j)
            for (int i = 0; i < d[0]; ++i) {
                const double x = coord(lo[0], hi[0], d[0], i);
                const double y = coord(lo[1], hi[1], d[1], j);
                const double z = coord(lo[2], hi[2], d[2], k);
                const double env = envelope(x, r0[0], sigma[0]) * envelope(y, r0[1], sigma[1]) *
                                   envelope(z, r0[2], sigma[2]);
                const double phase = k0[0] * x + k0[1] * y + k0[2] * z;
                const long f = i + (long)d[0] * (j + (long)d[1] * k);
                psi[2 * f] = env * cos(phase);
                psi[2 * f + 1] = env * sin(phase);
            }
    orc_normalize(psi, (size_t)d[0] * d[1] * d[2], cell_volume3(lo, hi, d));
}

/* ---- observables (core/src/observables.ixx) ----------------------------- */

static inline double obs_ratio(double num, double den) { return den > 0.0 ? num / den : 0.0; }

/* core/src/observables.ixx:29-38 */
double orc_mean_position_1d(const double* psi, int n, double xmin, double xmax) {
    double num = 0.0, den = 0.0;
    for (int i = 0; i < n; ++i) {
        const double w = psi[2 * i] * psi[2 * i] + psi[2 * i + 1] * psi[2 * i + 1];
        num += coord(xmin, xmax, n, i) * w;
        den += w;
    }
    return obs_ratio(num, den);
}

/* core/src/observables.ixx:67-90: each term normalised separately */
double orc_mean_energy_1d(const double* psi, int n, double xmin, double xmax, const double* v) {
    double num_v = 0.0, den_x = 0.0;
    for (int i = 0; i < n; ++i) {
        const double w = psi[2 * i] * psi[2 * i] + psi[2 * i + 1] * psi[2 * i + 1];
        num_v += v[i] * w;
        den_x += w;
    }
    double* phi = (double*)malloc(sizeof(double) * 2 * (size_t)n);
    memcpy(phi, psi, sizeof(double) * 2 * (size_t)n);
    orc_fft1d(phi, (size_t)n);
    double* k = (double*)malloc(sizeof(double) * (size_t)n);
    orc_wavenumbers(n, xmin, xmax, k);
    double num_t = 0.0, den_k = 0.0;
    for (int j = 0; j < n; ++j) {
        const double w = phi[2 * j] * phi[2 * j] + phi[2 * j + 1] * phi[2 * j + 1];
        num_t += 0.5 * k[j] * k[j] * w;
        den_k += w;
    }
    free(k);
    free(phi);
    return obs_ratio(num_v, den_x) + obs_ratio(num_t, den_k);
}

/* core/src/observables.ixx:92-119: Var(H) in the residual form
 * ||(H - <H>) psi||^2 / ||psi||^2 with H psi = ifft(0.5 k^2 fft(psi)) + V psi.
 * A length that is not a power of two (no reference counterpart: the
 * reference throws) takes the mixed-radix transform. */
double orc_energy_variance_1d(const double* psi, int n, double xmin, double xmax, const double* v) {
    double* h = (double*)malloc(sizeof(double) * 2 * (size_t)n);
    memcpy(h, psi, sizeof(double) * 2 * (size_t)n);
    const int pow2 = orc_fft1d(h, (size_t)n) == 0;
    if (!pow2) orc_fft_any(h, (size_t)n, -1);
    double* k = (double*)malloc(sizeof(double) * (size_t)n);
    orc_wavenumbers(n, xmin, xmax, k);
    for (int j = 0; j < n; ++j) {
        const double s = 0.5 * k[j] * k[j];
        h[2 * j] *= s;
        h[2 * j + 1] *= s;
    }
    if (pow2) {
        orc_ifft1d(h, (size_t)n);
    } else {
        /* the same conjugation identity as orc_ifft1d around the mixed-radix
         * forward transform */
        for (int i = 0; i < n; ++i) h[2 * i + 1] = -h[2 * i + 1];
        orc_fft_any(h, (size_t)n, -1);
        const double inv = 1.0 / (double)n;
        for (int i = 0; i < n; ++i) {
            h[2 * i] = inv * h[2 * i];
            h[2 * i + 1] = inv * -h[2 * i + 1];
        }
    }
    double num_h = 0.0, den = 0.0;
    for (int i = 0; i < n; ++i) {
        const double fr = psi[2 * i], fi = psi[2 * i + 1];
        h[2 * i] += v[i] * fr;
        h[2 * i + 1] += v[i] * fi;
        den += fr * fr + fi * fi;
        /* (conj(f) * h).real() */
        num_h += fr * h[2 * i] - (-fi) * h[2 * i + 1];
    }
    const double mean = obs_ratio(num_h, den);
    double num_var = 0.0;
    for (int i = 0; i < n; ++i) {
        const double dr = h[2 * i] - mean * psi[2 * i];
        const double di = h[2 * i + 1] - mean * psi[2 * i + 1];
        num_var += dr * dr + di * di;
    }
    free(k);
    free(h);
    return obs_ratio(num_var, den);
}

/* core/src/observables.ixx:121-131: not scale invariant, times dx */
double orc_probability_in_range_1d(const double* psi, int n, double xmin, double xmax, double a,
                                   double b) {
    double acc = 0.0;
    for (int i = 0; i < n; ++i) {
        const double x = coord(xmin, xmax, n, i);
        if (x >= a && x < b) acc += psi[2 * i] * psi[2 * i] + psi[2 * i + 1] * psi[2 * i + 1];
    }
    return acc * orc_spacing(xmin, xmax, n);
}

/* core/src/observables.ixx:134-153 */
void orc_mean_position_3d(const double* psi, const int d[3], const double lo[3],
                          const double hi[3], double out[3]) {
    double nx = 0.0, ny = 0.0, nz = 0.0, den = 0.0;
    for (int k = 0; k < d[2]; ++k)
        for (int j = 0; j < d[1]; ++j)
            for (int i = 0; i < d[0]; ++i) {
                const long f = i + (long)d[0] * (j + (long)d[1] * k);
                const double w = psi[2 * f] * psi[2 * f] + psi[2 * f + 1] * psi[2 * f + 1];
                nx += coord(lo[0], hi[0], d[0], i) * w;
                ny += coord(lo[1], hi[1], d[1], j) * w;
                nz += coord(lo[2], hi[2], d[2], k) * w;
                den += w;
            }
    out[0] = obs_ratio(nx, den);
    out[1] = obs_ratio(ny, den);
    out[2] = obs_ratio(nz, den);
}

static inline double obs_sigma(double second, double mean) {
    const double v = second - mean * mean;
    return sqrt(v > 0.0 ? v : 0.0);
}

/* core/src/observables.ixx:40-50 */
double orc_sigma_x_1d(const double* psi, int n, double xmin, double xmax) {
    double num = 0.0, den = 0.0;
    for (int i = 0; i < n; ++i) {
        const double x = coord(xmin, xmax, n, i);
        const double w = psi[2 * i] * psi[2 * i] + psi[2 * i + 1] * psi[2 * i + 1];
        num += x * x * w;
        den += w;
    }
    return obs_sigma(obs_ratio(num, den), orc_mean_position_1d(psi, n, xmin, xmax));
}

/* core/src/observables.ixx:52-65 (power-of-two n, like the reference) */
double orc_mean_momentum_1d(const double* psi, int n, double xmin, double xmax) {
    double* phi = (double*)malloc(sizeof(double) * 2 * (size_t)n);
    memcpy(phi, psi, sizeof(double) * 2 * (size_t)n);
    if (orc_fft1d(phi, (size_t)n) != 0) orc_fft_any(phi, (size_t)n, -1);
    double* k = (double*)malloc(sizeof(double) * (size_t)n);
    orc_wavenumbers(n, xmin, xmax, k);
    double num = 0.0, den = 0.0;
    for (int j = 0; j < n; ++j) {
        const double w = phi[2 * j] * phi[2 * j] + phi[2 * j + 1] * phi[2 * j + 1];
        num += k[j] * w;
        den += w;
    }
    free(k);
    free(phi);
    return obs_ratio(num, den);
}

/* core/src/observables.ixx:153-175 */
void orc_sigma_position_3d(const double* psi, const int d[3], const double lo[3],
                           const double hi[3], double out[3]) {
    double nx = 0.0, ny = 0.0, nz = 0.0, den = 0.0;
    for (int k = 0; k < d[2]; ++k)
        for (int j = 0; j < d[1]; ++j)
            for (int i = 0; i < d[0]; ++i) {
                const long f = i + (long)d[0] * (j + (long)d[1] * k);
                const double w = psi[2 * f] * psi[2 * f] + psi[2 * f + 1] * psi[2 * f + 1];
                const double x = coord(lo[0], hi[0], d[0], i);
                const double y = coord(lo[1], hi[1], d[1], j);
                const double z = coord(lo[2], hi[2], d[2], k);
                nx += x * x * w;
                ny += y * y * w;
                nz += z * z * w;
                den += w;
            }
    double m[3];
    orc_mean_position_3d(psi, d, lo, hi, m);
    out[0] = obs_sigma(obs_ratio(nx, den), m[0]);
    out[1] = obs_sigma(obs_ratio(ny, den), m[1]);
    out[2] = obs_sigma(obs_ratio(nz, den), m[2]);
}

/* core/src/observables.ixx:177-200 */
void orc_mean_momentum_3d(const double* psi, const int d[3], const double lo[3],
                          const double hi[3], double out[3]) {
    const long n = (long)d[0] * d[1] * d[2];
    double* phi = (double*)malloc(sizeof(double) * 2 * (size_t)n);
    memcpy(phi, psi, sizeof(double) * 2 * (size_t)n);
    if (orc_fft3d(phi, d[0], d[1], d[2]) != 0) orc_fft3d_any(phi, d[0], d[1], d[2], -1);
    double* kx = (double*)malloc(sizeof(double) * (size_t)d[0]);
    double* ky = (double*)malloc(sizeof(double) * (size_t)d[1]);
    double* kz = (double*)malloc(sizeof(double) * (size_t)d[2]);
    orc_wavenumbers(d[0], lo[0], hi[0], kx);
    orc_wavenumbers(d[1], lo[1], hi[1], ky);
    orc_wavenumbers(d[2], lo[2], hi[2], kz);
    double nx = 0.0, ny = 0.0, nz = 0.0, den = 0.0;
    for (int k = 0; k < d[2]; ++k)
        for (int j = 0; j < d[1]; ++j)
            for (int i = 0; i < d[0]; ++i) {
                const long f = i + (long)d[0] * (j + (long)d[1] * k);
                const double w = phi[2 * f] * phi[2 * f] + phi[2 * f + 1] * phi[2 * f + 1];
                nx += kx[i] * w;
                ny += ky[j] * w;
                nz += kz[k] * w;
                den += w;
            }
    free(kx);
    free(ky);
    free(kz);
    free(phi);
    out[0] = obs_ratio(nx, den);
    out[1] = obs_ratio(ny, den);
    out[2] = obs_ratio(nz, den);
}

/* core/src/observables.ixx:202-236 (power-of-two grids, like the reference) */
double orc_mean_energy_3d(const double* psi, const int d[3], const double lo[3],
                          const double hi[3], const double* v, double mass) {
    const long n = (long)d[0] * d[1] * d[2];
    double num_v = 0.0, den_x = 0.0;
    for (long i = 0; i < n; ++i) {
        const double w = psi[2 * i] * psi[2 * i] + psi[2 * i + 1] * psi[2 * i + 1];
        num_v += v[i] * w;
        den_x += w;
    }
    double* phi = (double*)malloc(sizeof(double) * 2 * (size_t)n);
    memcpy(phi, psi, sizeof(double) * 2 * (size_t)n);
    if (orc_fft3d(phi, d[0], d[1], d[2]) != 0) orc_fft3d_any(phi, d[0], d[1], d[2], -1);
    double* kx = (double*)malloc(sizeof(double) * (size_t)d[0]);
    double* ky = (double*)malloc(sizeof(double) * (size_t)d[1]);
    double* kz = (double*)malloc(sizeof(double) * (size_t)d[2]);
    orc_wavenumbers(d[0], lo[0], hi[0], kx);
    orc_wavenumbers(d[1], lo[1], hi[1], ky);
    orc_wavenumbers(d[2], lo[2], hi[2], kz);
    double num_t = 0.0, den_k = 0.0;
    for (int k = 0; k < d[2]; ++k)
        for (int j = 0; j < d[1]; ++j)
            for (int i = 0; i < d[0]; ++i) {
                const long f = i + (long)d[0] * (j + (long)d[1] * k);
                const double w = phi[2 * f] * phi[2 * f] + phi[2 * f + 1] * phi[2 * f + 1];
                const double kxx = kx[i], kyy = ky[j], kzz = kz[k];
                num_t += 0.5 * (kxx * kxx + kyy * kyy + kzz * kzz) / mass * w;
                den_k += w;
            }
    free(kx);
    free(ky);
    free(kz);
    free(phi);
    return obs_ratio(num_v, den_x) + obs_ratio(num_t, den_k);
}

/* No 3D counterpart in the reference (it has the 1D form only,
 * core/src/observables.ixx:121-131): the same half-open, not scale invariant
 * sum restricted to a slab a <= coord_axis < b, times the cell volume.
 * Pinned by the Gaussian closed form (erfc tail) in the oracle tests. */
double orc_probability_in_slab_3d(const double* psi, const int d[3], const double lo[3],
                                  const double hi[3], int axis, double a, double b) {
    double acc = 0.0;
    for (int k = 0; k < d[2]; ++k)
        for (int j = 0; j < d[1]; ++j)
            for (int i = 0; i < d[0]; ++i) {
                const int idx[3] = {i, j, k};
                const double x = coord(lo[axis], hi[axis], d[axis], idx[axis]);
                if (!(x >= a && x < b)) continue;
                const long f = i + (long)d[0] * (j + (long)d[1] * k);
                acc += psi[2 * f] * psi[2 * f] + psi[2 * f + 1] * psi[2 * f + 1];
            }
    return acc * cell_volume3(lo, hi, d);
}

/* No counterpart in the reference (extension of the slab above to a box):
 * the sum over cells with a[ax] <= coord_ax < b[ax] on every axis of extent
 * above 1 (axes of extent 1 are not restricted), times the cell volume.
 * Pinned in the oracle tests by the slab (one restricted axis), by the
 * product of the three slabs for a separable state, and by the Gaussian
 * closed form. */
double orc_probability_in_box_3d(const double* psi, const int d[3], const double lo[3],
                                 const double hi[3], const double a[3], const double b[3]) {
    double acc = 0.0;
    for (int k = 0; k < d[2]; ++k)
        for (int j = 0; j < d[1]; ++j)
            for (int i = 0; i < d[0]; ++i) {
                const int idx[3] = {i, j, k};
                int in = 1;
                for (int ax = 0; ax < 3; ++ax) {
                    if (d[ax] == 1) continue;
                    const double x = coord(lo[ax], hi[ax], d[ax], idx[ax]);
                    if (!(x >= a[ax] && x < b[ax])) in = 0;
                }
                if (!in) continue;
                const long f = i + (long)d[0] * (j + (long)d[1] * k);
                acc += psi[2 * f] * psi[2 * f] + psi[2 * f + 1] * psi[2 * f + 1];
            }
    return acc * cell_volume3(lo, hi, d);
}
