// builders.cpp -- host builders of the inputs a solve needs: potentials, the
// absorbing mask and Gaussian packets, in FP64 on the x-fastest grid.
//
// They restate the reference builders (core/src/potential.ixx,
// core/src/wavepacket.ixx) expression for expression and are compiled with
// -ffp-contract=off, so for the same arguments they return the same bits as
// the reference.  Grids follow core/src/grid.ixx: spacing (hi - lo) / n,
// coord(i) = lo + i * spacing.
#include <cmath>
#include <cstddef>
#include <vector>

#include "../../include/tdse.h"

namespace {

constexpr double kPi = 3.14159265358979323846;

inline double spacing(double lo, double hi, int n) { return (hi - lo) / n; }
inline double coord(double lo, double hi, int n, int i) { return lo + i * spacing(lo, hi, n); }
inline long long flat(const int n[3], int i, int j, int k) {
    return i + static_cast<long long>(n[0]) * (j + static_cast<long long>(n[1]) * k);
}
bool ok_grid(const int n[3], const double lo[3], const double hi[3]) {
    if (!n || !lo || !hi) return false;
    for (int a = 0; a < 3; ++a)
        if (n[a] < 1) return false;
    return true;
}

// core/src/potential.ixx:194-215
void mask_1d(double lo, double hi, int n, double width, double* m) {
    if (n == 1) {
        m[0] = 1.0;
        return;
    }
    for (int i = 0; i < n; ++i) {
        const double x = coord(lo, hi, n, i);
        const double d_lo = x - lo;
        const double d_hi = hi - x;
        const double d = d_lo < d_hi ? d_lo : d_hi;
        if (d >= width) {
            m[i] = 1.0;
            continue;
        }
        const double t = d / width;
        const double s = std::sin(0.5 * 3.14159265358979323846 * t);
        m[i] = s * s;
    }
}

double envelope(double u, double u0, double s) {
    return std::pow(2.0 * kPi * s * s, -0.25) * std::exp(-(u - u0) * (u - u0) / (4.0 * s * s));
}

}  // namespace

extern "C" {

// core/src/potential.ixx:16-23 (1D) and 34-48 (3D): 0.5 w^2 |r - c|^2
int tdse_build_harmonic(const int n[3], const double lo[3], const double hi[3], double omega,
                        const double c[3], double* v) {
    if (!ok_grid(n, lo, hi) || !c || !v) return TDSE_E_INVALID;
    for (int k = 0; k < n[2]; ++k)
        for (int j = 0; j < n[1]; ++j)
            for (int i = 0; i < n[0]; ++i) {
                const double dx = coord(lo[0], hi[0], n[0], i) - c[0];
                if (n[1] == 1 && n[2] == 1) {
                    v[i] = 0.5 * omega * omega * dx * dx;  // the 1D builder's grouping
                    continue;
                }
                const double dy = coord(lo[1], hi[1], n[1], j) - c[1];
                const double dz = coord(lo[2], hi[2], n[2], k) - c[2];
                v[flat(n, i, j, k)] = 0.5 * omega * omega * (dx * dx + dy * dy + dz * dz);
            }
    return TDSE_OK;
}

// core/src/potential.ixx:102-132: v0 on lo <= x < hi (half-open), along x
int tdse_build_barrier(const int n[3], const double lo[3], const double hi[3], double v0,
                       double x_lo, double x_hi, double* v) {
    if (!ok_grid(n, lo, hi) || !v) return TDSE_E_INVALID;
    for (int k = 0; k < n[2]; ++k)
        for (int j = 0; j < n[1]; ++j)
            for (int i = 0; i < n[0]; ++i) {
                const double x = coord(lo[0], hi[0], n[0], i);
                v[flat(n, i, j, k)] = (x >= x_lo && x < x_hi) ? v0 : 0.0;
            }
    return TDSE_OK;
}

// core/src/potential.ixx:134-154: -Z/r, the nucleus cell -Z * 2.3800774 / h
int tdse_build_regularized_coulomb(const int n[3], const double lo[3], const double hi[3],
                                   double z, const double c[3], double* v) {
    if (!ok_grid(n, lo, hi) || !c || !v) return TDSE_E_INVALID;
    const double h = spacing(lo[0], hi[0], n[0]);
    const double center = -z * 2.3800774 / h;
    for (int k = 0; k < n[2]; ++k)
        for (int j = 0; j < n[1]; ++j)
            for (int i = 0; i < n[0]; ++i) {
                const double dx = coord(lo[0], hi[0], n[0], i) - c[0];
                const double dy = coord(lo[1], hi[1], n[1], j) - c[1];
                const double dz = coord(lo[2], hi[2], n[2], k) - c[2];
                const double r = std::sqrt(dx * dx + dy * dy + dz * dz);
                v[flat(n, i, j, k)] = (r < 1e-6 * h) ? center : -z / r;
            }
    return TDSE_OK;
}

// The hydrogen-like workload's Coulomb with a floor: -Z / max(r, h/2).
// No reference builder has this form; it differs from the one above only at
// the nucleus cell (-2Z/h instead of -2.3800774 Z/h).
int tdse_build_coulomb_floor(const int n[3], const double lo[3], const double hi[3], double z,
                             const double c[3], double* v) {
    if (!ok_grid(n, lo, hi) || !c || !v) return TDSE_E_INVALID;
    const double h = spacing(lo[0], hi[0], n[0]);
    const double rmin = 0.5 * h;
    for (int k = 0; k < n[2]; ++k)
        for (int j = 0; j < n[1]; ++j)
            for (int i = 0; i < n[0]; ++i) {
                const double dx = coord(lo[0], hi[0], n[0], i) - c[0];
                const double dy = coord(lo[1], hi[1], n[1], j) - c[1];
                const double dz = coord(lo[2], hi[2], n[2], k) - c[2];
                const double r = std::sqrt(dx * dx + dy * dy + dz * dz);
                v[flat(n, i, j, k)] = -z / (r > rmin ? r : rmin);
            }
    return TDSE_OK;
}

// core/src/potential.ixx:25-33 (1D, n x 1 x 1) and 51-65 (3D): -Z / sqrt(r^2 + a^2)
int tdse_build_soft_coulomb(const int n[3], const double lo[3], const double hi[3], double z,
                            double a, const double c[3], double* v) {
    if (!ok_grid(n, lo, hi) || !c || !v) return TDSE_E_INVALID;
    for (int k = 0; k < n[2]; ++k)
        for (int j = 0; j < n[1]; ++j)
            for (int i = 0; i < n[0]; ++i) {
                const double dx = coord(lo[0], hi[0], n[0], i) - c[0];
                if (n[1] == 1 && n[2] == 1) {
                    v[i] = -z / std::sqrt(dx * dx + a * a);  // the 1D builder's form
                    continue;
                }
                const double dy = coord(lo[1], hi[1], n[1], j) - c[1];
                const double dz = coord(lo[2], hi[2], n[2], k) - c[2];
                v[flat(n, i, j, k)] = -z / std::sqrt(dx * dx + dy * dy + dz * dz + a * a);
            }
    return TDSE_OK;
}

// core/src/potential.ixx:194-233: separable sin^2 ramp, 1 on collapsed axes
int tdse_build_absorbing_mask(const int n[3], const double lo[3], const double hi[3],
                              double width, double* m) {
    if (!ok_grid(n, lo, hi) || !m) return TDSE_E_INVALID;
    double* mx = new double[n[0]];
    double* my = new double[n[1]];
    double* mz = new double[n[2]];
    mask_1d(lo[0], hi[0], n[0], width, mx);
    mask_1d(lo[1], hi[1], n[1], width, my);
    mask_1d(lo[2], hi[2], n[2], width, mz);
    for (int k = 0; k < n[2]; ++k)
        for (int j = 0; j < n[1]; ++j)
            for (int i = 0; i < n[0]; ++i) m[flat(n, i, j, k)] = mx[i] * my[j] * mz[k];
    delete[] mx;
    delete[] my;
    delete[] mz;
    return TDSE_OK;
}

// core/src/wavepacket.ixx:16-48: normalised Gaussian packet; sigma is the
// density standard deviation per axis.  A grid n x 1 x 1 uses the 1D form
// (core/src/wavepacket.ixx:16-25) and its normalisation (field.ixx:41-46).
int tdse_build_gaussian(const int n[3], const double lo[3], const double hi[3], const double r0[3],
                        const double sigma[3], const double k0[3], double* psi) {
    if (!ok_grid(n, lo, hi) || !r0 || !sigma || !k0 || !psi) return TDSE_E_INVALID;
    const long long total = static_cast<long long>(n[0]) * n[1] * n[2];
    if (n[1] == 1 && n[2] == 1) {
        const double amp = std::pow(2.0 * kPi * sigma[0] * sigma[0], -0.25);
        for (int i = 0; i < n[0]; ++i) {
            const double x = coord(lo[0], hi[0], n[0], i);
            const double env =
                amp * std::exp(-(x - r0[0]) * (x - r0[0]) / (4.0 * sigma[0] * sigma[0]));
            psi[2 * i] = env * std::cos(k0[0] * x);
            psi[2 * i + 1] = env * std::sin(k0[0] * x);
        }
        double acc = 0.0;
        for (int i = 0; i < n[0]; ++i) acc += psi[2 * i] * psi[2 * i] + psi[2 * i + 1] * psi[2 * i + 1];
        const double inv = 1.0 / std::sqrt(acc * spacing(lo[0], hi[0], n[0]));
        for (int i = 0; i < 2 * n[0]; ++i) psi[i] = inv * psi[i];
        return TDSE_OK;
    }
    // envelope() is a pure function of one coordinate: evaluate it once per
    // grid line index (same doubles as evaluating it per cell)
    std::vector<double> ex(n[0]), ey(n[1]), ez(n[2]);
    for (int i = 0; i < n[0]; ++i) ex[i] = envelope(coord(lo[0], hi[0], n[0], i), r0[0], sigma[0]);
    for (int j = 0; j < n[1]; ++j) ey[j] = envelope(coord(lo[1], hi[1], n[1], j), r0[1], sigma[1]);
    for (int k = 0; k < n[2]; ++k) ez[k] = envelope(coord(lo[2], hi[2], n[2], k), r0[2], sigma[2]);
    for (int k = 0; k < n[2]; ++k)
        for (int j = 0; j < n[1]; ++j)
            for (int i = 0; i < n[0]; ++i) {
                const double x = coord(lo[0], hi[0], n[0], i);
                const double y = coord(lo[1], hi[1], n[1], j);
                const double z = coord(lo[2], hi[2], n[2], k);
                const double env = ex[i] * ey[j] * ez[k];
                const double phase = k0[0] * x + k0[1] * y + k0[2] * z;
                const long long f = flat(n, i, j, k);
                psi[2 * f] = env * std::cos(phase);
                psi[2 * f + 1] = env * std::sin(phase);
            }
    // core/src/field.ixx:70-87
    double acc = 0.0;
    for (long long f = 0; f < total; ++f) acc += psi[2 * f] * psi[2 * f] + psi[2 * f + 1] * psi[2 * f + 1];
    const double n2 = acc * (spacing(lo[0], hi[0], n[0]) * spacing(lo[1], hi[1], n[1]) *
                             spacing(lo[2], hi[2], n[2]));
    if (n2 <= 0.0) return TDSE_OK;
    const double inv = 1.0 / std::sqrt(n2);
    for (long long f = 0; f < 2 * total; ++f) psi[f] = inv * psi[f];
    return TDSE_OK;
}

// The reference propagator's FP64 tables (core/src/propagator.ixx:23-40 for
// an n x 1 x 1 grid, 72-101 otherwise): half_v[f] = cis(-0.5 V[f] dt) and the
// full-grid kinetic[f] = cis(-0.5 (kx^2 + ky^2 + kz^2) / mass dt) with the
// fftfreq wavenumbers of core/src/spectral.ixx:14-23, cos and sin as two libm
// calls, interleaved complex128 out.  Either output may be null.
int tdse_build_phase_tables(const int n[3], const double lo[3], const double hi[3], const double* v,
                            double dt, double mass, double* half_v, double* kinetic) {
    if (!ok_grid(n, lo, hi) || (half_v && !v) || !(mass != 0.0)) return TDSE_E_INVALID;
    const long long total = static_cast<long long>(n[0]) * n[1] * n[2];
    if (half_v)
        for (long long f = 0; f < total; ++f) {
            const double th = -0.5 * v[f] * dt;
            half_v[2 * f] = std::cos(th);
            half_v[2 * f + 1] = std::sin(th);
        }
    if (!kinetic) return TDSE_OK;
    std::vector<double> k[3];
    for (int a = 0; a < 3; ++a) {
        const double dk = 2.0 * kPi / (hi[a] - lo[a]);
        k[a].resize(static_cast<std::size_t>(n[a]));
        for (int j = 0; j < n[a]; ++j) k[a][j] = dk * ((2 * j < n[a]) ? j : j - n[a]);
    }
    if (n[1] == 1 && n[2] == 1) {
        // SplitOperator1D: no mass
        for (int j = 0; j < n[0]; ++j) {
            const double th = -0.5 * k[0][j] * k[0][j] * dt;
            kinetic[2 * j] = std::cos(th);
            kinetic[2 * j + 1] = std::sin(th);
        }
        return TDSE_OK;
    }
    for (int kk = 0; kk < n[2]; ++kk)
        for (int j = 0; j < n[1]; ++j)
            for (int i = 0; i < n[0]; ++i) {
                const double kxx = k[0][i], kyy = k[1][j], kzz = k[2][kk];
                const double th = -0.5 * (kxx * kxx + kyy * kyy + kzz * kzz) / mass * dt;
                const long long f = flat(n, i, j, kk);
                kinetic[2 * f] = std::cos(th);
                kinetic[2 * f + 1] = std::sin(th);
            }
    return TDSE_OK;
}

}  // extern "C"
