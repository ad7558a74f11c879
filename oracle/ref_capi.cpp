// ref_capi.cpp -- extern "C" shim over the reference core, compiled from the
// reference's own module sources (oracle/Makefile, target "ref").
//
// TEST INFRASTRUCTURE ONLY.  It lets the tests compare the C restatement in
// tdse_oracle.c bit for bit with the reference implementation, and lets the
// benchmark time the reference's own CPU propagator (its thread pool included)
// as the CPU baseline.  Nothing in the product links it.
//
// Complex arrays cross as interleaved (re, im) doubles; the shim copies them
// into ses::Field1D / ses::Field3D (std::complex<double>, same layout).

#include <complex>
#include <cstddef>
#include <cstring>
#include <stdexcept>
#include <vector>

import ses.grid;
import ses.vec;
import ses.field;
import ses.spectral;
import ses.fft;
import ses.propagator;
import ses.potential;
import ses.wavepacket;
import ses.observables;
import ses.parallel;
import ses.imaginary_time;
import ses.drive;
import ses.rotation;
import ses.magnetic;
import ses.lattice2d;

namespace {

using Cd = std::complex<double>;

ses::Grid3D grid3(const int d[3], const double lo[3], const double hi[3]) {
    return ses::Grid3D{ses::Grid1D{lo[0], hi[0], d[0]}, ses::Grid1D{lo[1], hi[1], d[1]},
                       ses::Grid1D{lo[2], hi[2], d[2]}};
}

void load(std::vector<Cd>& dst, const double* src) {
    std::memcpy(static_cast<void*>(dst.data()), src, dst.size() * sizeof(Cd));
}
void store(double* dst, const std::vector<Cd>& src) {
    std::memcpy(dst, static_cast<const void*>(src.data()), src.size() * sizeof(Cd));
}

}  // namespace

extern "C" {

int ref_parallel_workers() { return ses::parallel_workers(); }

void ref_wavenumbers(int n, double xmin, double xmax, double* k) {
    const std::vector<double> v = ses::wavenumbers(ses::Grid1D{xmin, xmax, n});
    std::memcpy(k, v.data(), v.size() * sizeof(double));
}

int ref_fft1d(double* a, std::size_t n) {
    std::vector<Cd> v(n);
    load(v, a);
    try {
        ses::fft(v);
    } catch (const std::invalid_argument&) {
        return -1;
    }
    store(a, v);
    return 0;
}

int ref_ifft1d(double* a, std::size_t n) {
    std::vector<Cd> v(n);
    load(v, a);
    try {
        ses::ifft(v);
    } catch (const std::invalid_argument&) {
        return -1;
    }
    store(a, v);
    return 0;
}

int ref_fft3d(double* a, const int d[3]) {
    const double lo[3] = {0, 0, 0}, hi[3] = {1, 1, 1};
    ses::Field3D f{grid3(d, lo, hi)};
    load(f.data(), a);
    try {
        ses::fft(f);
    } catch (const std::invalid_argument&) {
        return -1;
    }
    store(a, f.data());
    return 0;
}

int ref_ifft3d(double* a, const int d[3]) {
    const double lo[3] = {0, 0, 0}, hi[3] = {1, 1, 1};
    ses::Field3D f{grid3(d, lo, hi)};
    load(f.data(), a);
    try {
        ses::ifft(f);
    } catch (const std::invalid_argument&) {
        return -1;
    }
    store(a, f.data());
    return 0;
}

// SplitOperator1D: evolve psi in place by nsteps (core/src/propagator.ixx:21-64).
int ref_so1d_evolve(const double* v, int n, double xmin, double xmax, double dt, double* psi,
                    int nsteps, const double* mask) {
    const ses::Grid1D g{xmin, xmax, n};
    std::vector<double> pot(v, v + n);
    ses::SplitOperator1D op{g, pot, dt};
    ses::Field1D f{g};
    load(f.data(), psi);
    try {
        if (mask == nullptr) {
            op.step(f, nsteps);
        } else {
            for (int s = 0; s < nsteps; ++s) {
                op.step(f, 1);
                for (int i = 0; i < n; ++i) f[i] *= mask[i];
            }
        }
    } catch (const std::invalid_argument&) {
        return -1;
    }
    store(psi, f.data());
    return 0;
}

// SplitOperator3D tables through the reference accessors (propagator.ixx:107-108).
void ref_so3d_tables(const double* v, const int d[3], const double lo[3], const double hi[3],
                     double dt, double mass, double* half_v, double* kinetic) {
    const ses::Grid3D g = grid3(d, lo, hi);
    std::vector<double> pot(v, v + g.size());
    const ses::SplitOperator3D op{g, pot, dt, mass};
    store(half_v, op.half_potential_phase());
    store(kinetic, op.kinetic_phase());
}

// SplitOperator3D::step, with the caller-side mask after each step when given.
int ref_so3d_evolve(const double* v, const int d[3], const double lo[3], const double hi[3],
                    double dt, double mass, double* psi, int nsteps, const double* mask) {
    const ses::Grid3D g = grid3(d, lo, hi);
    std::vector<double> pot(v, v + g.size());
    const ses::SplitOperator3D op{g, pot, dt, mass};
    ses::Field3D f{g};
    load(f.data(), psi);
    try {
        if (mask == nullptr) {
            op.step(f, nsteps);
        } else {
            const std::size_t n = f.data().size();
            for (int s = 0; s < nsteps; ++s) {
                op.step(f, 1);
                for (std::size_t i = 0; i < n; ++i) f.data()[i] *= mask[i];
            }
        }
    } catch (const std::invalid_argument&) {
        return -1;
    }
    store(psi, f.data());
    return 0;
}

// ses::driven_step with a SplitOperator3D (core/src/drive.ixx:52-66).
int ref_driven_step(const double* v, const int d[3], const double lo[3], const double hi[3],
                    double dt, double mass, double* psi, const double axis[3], double amplitude,
                    double omega, double t0, int nsteps) {
    const ses::Grid3D g = grid3(d, lo, hi);
    std::vector<double> pot(v, v + g.size());
    const ses::SplitOperator3D op{g, pot, dt, mass};
    ses::Field3D f{g};
    load(f.data(), psi);
    try {
        ses::driven_step(f, op, ses::DipoleDrive{ses::Vec3d{axis[0], axis[1], axis[2]}, amplitude, omega},
                         t0, nsteps);
    } catch (const std::invalid_argument&) {
        return -1;
    }
    store(psi, f.data());
    return 0;
}

// ses::rotate_axis (core/src/rotation.ixx:73-84).
int ref_rotate_axis(double* psi, const int d[3], const double lo[3], const double hi[3], int axis,
                    double theta) {
    ses::Field3D f{grid3(d, lo, hi)};
    load(f.data(), psi);
    try {
        ses::rotate_axis(f, axis, theta);
    } catch (const std::invalid_argument&) {
        return -1;
    }
    store(psi, f.data());
    return 0;
}

// MagneticPropagator3D::step and effective_potential (core/src/magnetic.ixx:19-80).
int ref_magnetic_evolve(const double* v, const int d[3], const double lo[3], const double hi[3],
                        double dt, double bfield, int axis, double* psi, int nsteps,
                        double* veff_out) {
    const ses::Grid3D g = grid3(d, lo, hi);
    std::vector<double> pot(v, v + g.size());
    const ses::MagneticPropagator3D op{g, pot, dt, bfield, axis};
    if (veff_out)
        std::memcpy(veff_out, op.effective_potential().data(), g.size() * sizeof(double));
    ses::Field3D f{g};
    load(f.data(), psi);
    try {
        op.step(f, nsteps);
    } catch (const std::invalid_argument&) {
        return -1;
    }
    store(psi, f.data());
    return 0;
}

// PeierlsLattice2D (core/src/lattice2d.ixx:31-275).  links: 0 none,
// 1 uniform field a, 2 solenoid (flux a at (xs, ys), cut_up).  Writes the
// x-links (link_x(i, j), interleaved) when links_out is given and returns
// the energy of the final state in *energy when given.
int ref_lat2d_evolve(const double* v, const int d[3], const double lo[3], const double hi[3],
                     double dt, double* psi, int nsteps, int relax, int links, double a, double xs,
                     double ys, int cut_up, double* links_out, double* energy) {
    const ses::Grid3D g = grid3(d, lo, hi);
    std::vector<double> pot(v, v + g.size());
    ses::PeierlsLattice2D prop{g, pot, dt};
    if (links == 1) prop.set_uniform_field(a);
    if (links == 2) prop.set_solenoid(a, xs, ys, cut_up != 0);
    ses::Field3D f{g};
    load(f.data(), psi);
    if (relax) prop.relax(f, nsteps);
    else prop.step(f, nsteps);
    store(psi, f.data());
    if (links_out) {
        for (int j = 0; j < d[1]; ++j)
            for (int i = 0; i < d[0]; ++i) {
                const Cd u = prop.link_x(i, j);
                links_out[2 * (static_cast<std::size_t>(j) * d[0] + i)] = u.real();
                links_out[2 * (static_cast<std::size_t>(j) * d[0] + i) + 1] = u.imag();
            }
    }
    if (energy) *energy = prop.energy(f);
    return 0;
}

// Widths and momenta (core/src/observables.ixx:40-65, 153-200).
double ref_sigma_x_1d(const double* psi, int n, double xmin, double xmax) {
    ses::Field1D f{ses::Grid1D{xmin, xmax, n}};
    load(f.data(), psi);
    return ses::sigma_x(f);
}
double ref_mean_momentum_1d(const double* psi, int n, double xmin, double xmax) {
    ses::Field1D f{ses::Grid1D{xmin, xmax, n}};
    load(f.data(), psi);
    return ses::mean_momentum(f);
}
void ref_sigma_position_3d(const double* psi, const int d[3], const double lo[3],
                           const double hi[3], double* out) {
    ses::Field3D f{grid3(d, lo, hi)};
    load(f.data(), psi);
    const ses::Vec3d s = ses::sigma_position(f);
    out[0] = s.x;
    out[1] = s.y;
    out[2] = s.z;
}
void ref_mean_momentum_3d(const double* psi, const int d[3], const double lo[3],
                          const double hi[3], double* out) {
    ses::Field3D f{grid3(d, lo, hi)};
    load(f.data(), psi);
    const ses::Vec3d s = ses::mean_momentum(f);
    out[0] = s.x;
    out[1] = s.y;
    out[2] = s.z;
}

// ImaginaryTimePropagator1D::relax (core/src/imaginary_time.ixx:21-63).
int ref_it1d_relax(const double* v, int n, double xmin, double xmax, double dtau, double* psi,
                   int nsteps) {
    const ses::Grid1D g{xmin, xmax, n};
    std::vector<double> pot(v, v + n);
    const ses::ImaginaryTimePropagator1D op{g, pot, dtau};
    ses::Field1D f{g};
    load(f.data(), psi);
    try {
        op.relax(f, nsteps);
    } catch (const std::invalid_argument&) {
        return -1;
    }
    store(psi, f.data());
    return 0;
}

// ImaginaryTimePropagator3D weights through its accessors (imaginary_time.ixx:96-98).
void ref_it3d_tables(const double* v, const int d[3], const double lo[3], const double hi[3],
                     double dtau, double mass, double* half_w, double* kin_w) {
    const ses::Grid3D g = grid3(d, lo, hi);
    std::vector<double> pot(v, v + g.size());
    const ses::ImaginaryTimePropagator3D op{g, pot, dtau, mass};
    std::memcpy(half_w, op.half_potential_weight().data(), g.size() * sizeof(double));
    std::memcpy(kin_w, op.kinetic_weight().data(), g.size() * sizeof(double));
}

// ImaginaryTimePropagator3D::relax, or relax_deflated against `lower` states
// (imaginary_time.ixx:100-134).
int ref_it3d_relax(const double* v, const int d[3], const double lo[3], const double hi[3],
                   double dtau, double mass, double* psi, const double* const* lower, int nlower,
                   int nsteps) {
    const ses::Grid3D g = grid3(d, lo, hi);
    std::vector<double> pot(v, v + g.size());
    const ses::ImaginaryTimePropagator3D op{g, pot, dtau, mass};
    ses::Field3D f{g};
    load(f.data(), psi);
    std::vector<ses::Field3D> low;
    low.reserve(static_cast<std::size_t>(nlower));
    std::vector<const ses::Field3D*> ptrs;
    for (int j = 0; j < nlower; ++j) {
        low.emplace_back(g);
        load(low.back().data(), lower[j]);
    }
    for (const ses::Field3D& q : low) ptrs.push_back(&q);
    try {
        if (nlower == 0) op.relax(f, nsteps);
        else op.relax_deflated(f, ptrs, nsteps);
    } catch (const std::invalid_argument&) {
        return -1;
    }
    store(psi, f.data());
    return 0;
}

// core/src/field.ixx:88-99
void ref_inner_product(const double* a, const double* b, const int d[3], const double lo[3],
                       const double hi[3], double* out) {
    const ses::Grid3D g = grid3(d, lo, hi);
    ses::Field3D fa{g}, fb{g};
    load(fa.data(), a);
    load(fb.data(), b);
    const Cd c = ses::inner_product(fa, fb);
    out[0] = c.real();
    out[1] = c.imag();
}

void ref_soft_coulomb_1d(double xmin, double xmax, int n, double z, double a, double x0,
                         double* v) {
    const std::vector<double> r = ses::soft_coulomb_potential(ses::Grid1D{xmin, xmax, n}, z, a, x0);
    std::memcpy(v, r.data(), r.size() * sizeof(double));
}
void ref_soft_coulomb_3d(const int d[3], const double lo[3], const double hi[3], double z, double a,
                         const double c[3], double* v) {
    const std::vector<double> r =
        ses::soft_coulomb_potential(grid3(d, lo, hi), z, a, ses::Vec3d{c[0], c[1], c[2]});
    std::memcpy(v, r.data(), r.size() * sizeof(double));
}

// Builders (core/src/potential.ixx, core/src/wavepacket.ixx).
void ref_harmonic_1d(double xmin, double xmax, int n, double omega, double x0, double* v) {
    const std::vector<double> r = ses::harmonic_potential(ses::Grid1D{xmin, xmax, n}, omega, x0);
    std::memcpy(v, r.data(), r.size() * sizeof(double));
}
void ref_harmonic_3d(const int d[3], const double lo[3], const double hi[3], double omega,
                     const double c[3], double* v) {
    const std::vector<double> r =
        ses::harmonic_potential(grid3(d, lo, hi), omega, ses::Vec3d{c[0], c[1], c[2]});
    std::memcpy(v, r.data(), r.size() * sizeof(double));
}
void ref_barrier_1d(double xmin, double xmax, int n, double v0, double x_lo, double x_hi,
                    double* v) {
    const std::vector<double> r = ses::barrier_potential(ses::Grid1D{xmin, xmax, n}, v0, x_lo, x_hi);
    std::memcpy(v, r.data(), r.size() * sizeof(double));
}
void ref_barrier_3d(const int d[3], const double lo[3], const double hi[3], double v0,
                    double x_lo, double x_hi, double* v) {
    const std::vector<double> r = ses::barrier_potential(grid3(d, lo, hi), v0, x_lo, x_hi);
    std::memcpy(v, r.data(), r.size() * sizeof(double));
}
void ref_regularized_coulomb_3d(const int d[3], const double lo[3], const double hi[3],
                                double z, const double c[3], double* v) {
    const std::vector<double> r =
        ses::regularized_coulomb_potential(grid3(d, lo, hi), z, ses::Vec3d{c[0], c[1], c[2]});
    std::memcpy(v, r.data(), r.size() * sizeof(double));
}
void ref_absorbing_mask_1d(double xmin, double xmax, int n, double width, double* m) {
    const std::vector<double> r = ses::absorbing_mask(ses::Grid1D{xmin, xmax, n}, width);
    std::memcpy(m, r.data(), r.size() * sizeof(double));
}
void ref_absorbing_mask_3d(const int d[3], const double lo[3], const double hi[3], double width,
                           double* m) {
    const std::vector<double> r = ses::absorbing_mask(grid3(d, lo, hi), width);
    std::memcpy(m, r.data(), r.size() * sizeof(double));
}
void ref_gaussian_1d(double xmin, double xmax, int n, double x0, double sigma, double k0,
                     double* psi) {
    const ses::Field1D f = ses::gaussian_wavepacket(ses::Grid1D{xmin, xmax, n}, x0, sigma, k0);
    store(psi, f.data());
}
void ref_gaussian_3d(const int d[3], const double lo[3], const double hi[3], const double r0[3],
                     const double sigma[3], const double k0[3], double* psi) {
    const ses::Field3D f = ses::gaussian_wavepacket(grid3(d, lo, hi), ses::Vec3d{r0[0], r0[1], r0[2]},
                                                    ses::Vec3d{sigma[0], sigma[1], sigma[2]},
                                                    ses::Vec3d{k0[0], k0[1], k0[2]});
    store(psi, f.data());
}

// Reductions and observables (core/src/field.ixx, core/src/observables.ixx).
double ref_norm_sq_1d(const double* psi, int n, double xmin, double xmax) {
    ses::Field1D f{ses::Grid1D{xmin, xmax, n}};
    load(f.data(), psi);
    return ses::norm_sq(f);
}
double ref_norm_sq_3d(const double* psi, const int d[3], const double lo[3], const double hi[3]) {
    ses::Field3D f{grid3(d, lo, hi)};
    load(f.data(), psi);
    return ses::norm_sq(f);
}
double ref_mean_position_1d(const double* psi, int n, double xmin, double xmax) {
    ses::Field1D f{ses::Grid1D{xmin, xmax, n}};
    load(f.data(), psi);
    return ses::mean_position(f);
}
double ref_mean_energy_1d(const double* psi, int n, double xmin, double xmax, const double* v) {
    ses::Field1D f{ses::Grid1D{xmin, xmax, n}};
    load(f.data(), psi);
    return ses::mean_energy(f, std::vector<double>(v, v + n));
}
double ref_energy_variance_1d(const double* psi, int n, double xmin, double xmax,
                              const double* v) {
    ses::Field1D f{ses::Grid1D{xmin, xmax, n}};
    load(f.data(), psi);
    return ses::energy_variance(f, std::vector<double>(v, v + n));
}
double ref_probability_in_range_1d(const double* psi, int n, double xmin, double xmax, double a,
                                   double b) {
    ses::Field1D f{ses::Grid1D{xmin, xmax, n}};
    load(f.data(), psi);
    return ses::probability_in_range(f, a, b);
}
void ref_mean_position_3d(const double* psi, const int d[3], const double lo[3],
                          const double hi[3], double out[3]) {
    ses::Field3D f{grid3(d, lo, hi)};
    load(f.data(), psi);
    const ses::Vec3d m = ses::mean_position(f);
    out[0] = m.x;
    out[1] = m.y;
    out[2] = m.z;
}
double ref_mean_energy_3d(const double* psi, const int d[3], const double lo[3],
                          const double hi[3], const double* v, double mass) {
    ses::Field3D f{grid3(d, lo, hi)};
    load(f.data(), psi);
    return ses::mean_energy(f, std::vector<double>(v, v + f.data().size()), mass);
}

}  // extern "C"
