/*
 * tdse_oracle.h -- CPU FP64 restatement of the reference split-operator path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the parity checker: only the
 * tests, the smoke check and the CPU-baseline leg of the benchmark load it.
 * The product (the HIP propagator behind include/tdse.h) never calls it.
 *
 * Layout conventions follow the reference core:
 *   - complex arrays are interleaved (re, im) doubles, x-fastest
 *     (core/src/grid.ixx:24, core/src/field.ixx:11-66);
 *   - grids are periodic with spacing (xmax - xmin) / n
 *     (core/src/grid.ixx:15-17).
 *
 * Every function names the reference lines it restates.  Functions with no
 * reference counterpart (mixed-radix FFT, 3D probability in a slab or a box,
 * the -1/max(r, h/2) Coulomb builder) say so and how they are pinned.
 */
#ifndef TDSE_ORACLE_H
#define TDSE_ORACLE_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Grid and spectral helpers (core/src/grid.ixx:10-30, core/src/spectral.ixx:14-23). */
double orc_spacing(double xmin, double xmax, int n);
void orc_wavenumbers(int n, double xmin, double xmax, double* k);

/* Radix-2 DIT FFT, forward unnormalised (core/src/fft.ixx:29-98).
 * Returns 0, or -1 for a size that is not a power of two (the reference
 * throws std::invalid_argument there, core/src/fft.ixx:42-44). */
void orc_fft_twiddles(size_t n, double* w);
int orc_fft_line(double* a, size_t n, const double* w);
int orc_fft1d(double* a, size_t n);
int orc_ifft1d(double* a, size_t n);
int orc_fft3d(double* a, int nx, int ny, int nz);
int orc_ifft3d(double* a, int nx, int ny, int nz);

/* Mixed-radix FFT for n = 2^a 3^b 5^c ... (no reference counterpart: the
 * reference throws for non powers of two).  Pinned by the brute-force DFT
 * below and by numpy.fft fixtures. */
int orc_fft_any(double* a, size_t n, int sign);
void orc_dft_brute(const double* in, double* out, size_t n, int sign);
int orc_fft3d_any(double* a, int nx, int ny, int nz, int sign);

/* Split-operator tables and Strang step (core/src/propagator.ixx:21-134). */
void orc_so1d_tables(const double* v, int n, double xmin, double xmax, double dt,
                     double* half_v, double* kinetic);
int orc_so1d_step(double* psi, int n, const double* half_v, const double* kinetic,
                  int nsteps);
void orc_so3d_tables(const double* v, const int dims[3], const double lo[3],
                     const double hi[3], double dt, double mass, double* half_v,
                     double* kinetic);
int orc_so3d_step(double* psi, const int dims[3], const double* half_v,
                  const double* kinetic, int nsteps);
/* Same step with the mixed-radix FFT (non power-of-two extension). */
int orc_so3d_step_any(double* psi, const int dims[3], const double* half_v,
                      const double* kinetic, int nsteps);
/* Caller-side absorbing mask after every step (scenario/src/line1d_director.ixx:202-211). */
void orc_apply_mask(double* psi, const double* mask, size_t n);

/* Field reductions (core/src/field.ixx:33-85). */
double orc_norm_sq(const double* psi, size_t n, double cell_volume);
void orc_normalize(double* psi, size_t n, double cell_volume);

/* Potential / mask / packet builders (core/src/potential.ixx, wavepacket.ixx). */
void orc_harmonic_1d(double xmin, double xmax, int n, double omega, double x0, double* v);
void orc_harmonic_3d(const int dims[3], const double lo[3], const double hi[3],
                     double omega, const double c[3], double* v);
void orc_barrier_1d(double xmin, double xmax, int n, double v0, double x_lo, double x_hi,
                    double* v);
void orc_barrier_3d(const int dims[3], const double lo[3], const double hi[3], double v0,
                    double x_lo, double x_hi, double* v);
void orc_regularized_coulomb_3d(const int dims[3], const double lo[3], const double hi[3],
                                double z, const double c[3], double* v);
void orc_coulomb_floor_3d(const int dims[3], const double lo[3], const double hi[3],
                          double z, const double c[3], double* v);
void orc_absorbing_mask_1d(double xmin, double xmax, int n, double width, double* m);
void orc_absorbing_mask_3d(const int dims[3], const double lo[3], const double hi[3],
                           double width, double* m);
void orc_gaussian_1d(double xmin, double xmax, int n, double x0, double sigma, double k0,
                     double* psi);
void orc_gaussian_3d(const int dims[3], const double lo[3], const double hi[3],
                     const double r0[3], const double sigma[3], const double k0[3],
                     double* psi);

/* Dipole drive (core/src/drive.ixx). */
void orc_dipole_halfkick(double* psi, const int dims[3], const double lo[3], const double hi[3],
                         const double axis[3], double amplitude, double omega, double t,
                         double dt);
int orc_driven_step(double* psi, const int dims[3], const double lo[3], const double hi[3],
                    const double* half_v, const double* kinetic, const double axis[3],
                    double amplitude, double omega, double t0, double dt, int nsteps,
                    int any_size);

/* Uniform magnetic field (core/src/magnetic.ixx, core/src/rotation.ixx). */
int orc_rotate_axis(double* psi, const int dims[3], const double lo[3], const double hi[3],
                    int axis, double theta);
void orc_magnetic_veff(const double* v, const int dims[3], const double lo[3],
                       const double hi[3], double bfield, int axis, double* out);
int orc_magnetic_step(double* psi, const int dims[3], const double lo[3], const double hi[3],
                      const double* half_v, const double* kinetic, double dt, double bfield,
                      int axis, int nsteps);

/* Peierls lattice (core/src/lattice2d.ixx). */
void orc_lat2d_tables(const double* v, int nx, int ny, double hx, double hy, double dt,
                      double* consts, double* half_v, double* relax_half_v);
void orc_lat2d_solenoid(double* link_x, int nx, int ny, double xmin, double xmax, double ymin,
                        double ymax, double phi, double xs, double ys, int cut_up);
void orc_lat2d_uniform_field(double* link_x, int nx, int ny, double xmin, double xmax,
                             double ymin, double ymax, double b);
int orc_lat2d_evolve(double* psi, int nx, int ny, const double* consts, const double* half_v,
                     const double* relax_half_v, const double* link_x, int relax, double dv,
                     int nsteps);
double orc_lat2d_energy(const double* psi, int nx, int ny, const double* v, const double* consts,
                        const double* link_x);

/* Imaginary-time relaxation (core/src/imaginary_time.ixx). */
void orc_it1d_tables(const double* v, int n, double xmin, double xmax, double dtau,
                     double* half_w, double* kin_w);
void orc_it3d_tables(const double* v, const int dims[3], const double lo[3],
                     const double hi[3], double dtau, double mass, double* half_w,
                     double* kin_w);
int orc_it1d_relax(double* psi, int n, double dx, const double* half_w, const double* kin_w,
                   int nsteps, double* last_norm);
int orc_it3d_relax(double* psi, const int dims[3], double dv, const double* half_w,
                   const double* kin_w, const double* const* lower, int nlower, int nsteps,
                   int any_size, double* last_norm);
void orc_inner_product(const double* a, const double* b, size_t n, double dv, double* out);
void orc_soft_coulomb_1d(double xmin, double xmax, int n, double z, double a, double x0,
                         double* v);
void orc_soft_coulomb_3d(const int dims[3], const double lo[3], const double hi[3], double z,
                         double a, const double c[3], double* v);

/* Observables (core/src/observables.ixx). */
double orc_mean_position_1d(const double* psi, int n, double xmin, double xmax);
double orc_energy_variance_1d(const double* psi, int n, double xmin, double xmax,
                              const double* v);
double orc_mean_energy_1d(const double* psi, int n, double xmin, double xmax,
                          const double* v);
double orc_probability_in_range_1d(const double* psi, int n, double xmin, double xmax,
                                   double a, double b);
void orc_mean_position_3d(const double* psi, const int dims[3], const double lo[3],
                          const double hi[3], double out[3]);
double orc_mean_energy_3d(const double* psi, const int dims[3], const double lo[3],
                          const double hi[3], const double* v, double mass);
double orc_sigma_x_1d(const double* psi, int n, double xmin, double xmax);
double orc_mean_momentum_1d(const double* psi, int n, double xmin, double xmax);
void orc_sigma_position_3d(const double* psi, const int dims[3], const double lo[3],
                           const double hi[3], double out[3]);
void orc_mean_momentum_3d(const double* psi, const int dims[3], const double lo[3],
                          const double hi[3], double out[3]);
double orc_probability_in_slab_3d(const double* psi, const int dims[3], const double lo[3],
                                  const double hi[3], int axis, double a, double b);
double orc_probability_in_box_3d(const double* psi, const int dims[3], const double lo[3],
                                 const double hi[3], const double a[3], const double b[3]);

#ifdef __cplusplus
}
#endif

#endif /* TDSE_ORACLE_H */
