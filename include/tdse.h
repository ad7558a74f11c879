/*
 * tdse.h -- C ABI of the MI355X split-operator TDSE propagator.
 *
 * Drop-in for the CPU FP64 propagator of the reference core:
 *   ses::SplitOperator1D(grid, V, dt)             core/src/propagator.ixx:23-40
 *   ses::SplitOperator3D(grid, V, dt, mass)       core/src/propagator.ixx:72-101
 *   step(Field&, int nsteps)                      core/src/propagator.ixx:44-53, 110-119
 *   WavepacketSimulation::advance(int)            core/src/simulation.ixx:40-43
 * plus the caller-side absorbing mask applied after every step
 * (scenario/src/line1d_director.ixx:202-211) and the observables of
 * core/src/field.ixx:33-85 and core/src/observables.ixx.
 *
 * Conventions (identical to the reference):
 *   - grids are periodic, spacing (hi - lo) / n, x-fastest flat index
 *     i + nx*(j + ny*k)  (core/src/grid.ixx:10-30);
 *   - a 2D grid is nx x ny x 1, a 1D grid is n x 1 x 1 (any axis of extent 1
 *     is dropped), as in the reference's Grid3D usage;
 *   - host states are interleaved complex128 (re, im) doubles, the layout of
 *     std::vector<std::complex<double>> (core/src/field.ixx:11-66).
 *
 * Differences from the reference interface, all in the direction of a C ABI:
 *   - every call returns a status (0 = TDSE_OK, < 0 = TDSE_E_*) instead of
 *     throwing std::invalid_argument (core/src/fft.ixx:42-44, 110-112);
 *   - the state stays resident on the GPU between calls: upload, step and
 *     download are separate calls (the reference GPU engine's
 *     upload_state/step/readback, solver/src/vk_engine.ixx:1083-1085, 2136-2157);
 *   - tdse_step only enqueues work on the plan's stream; tdse_plan_sync and
 *     tdse_state_download wait for it.
 *
 * Threading: one plan is used by one host thread at a time; distinct plans
 * may be driven from distinct threads.  Line lengths supported per axis:
 * 2..1024 powers of two and 96, 160, 192; a 1D grid also takes N = N1*N2
 * with N1, N2 in that set (e.g. 65536).
 */
#ifndef TDSE_H
#define TDSE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TDSE_OK 0
#define TDSE_E_INVALID (-1)     /* bad argument (null pointer, n < 1, dt not finite, ...) */
#define TDSE_E_UNSUPPORTED (-2) /* line length without a compiled FFT plan */
#define TDSE_E_OOM (-3)         /* device allocation failed or HBM budget exceeded */
#define TDSE_E_DEVICE (-4)      /* HIP runtime error; the plan's context is unusable */
#define TDSE_E_TIMEOUT (-5)     /* a bounded wait expired */
#define TDSE_E_NO_DEVICE (-6)   /* no HIP device with that ordinal */
#define TDSE_E_STATE (-7)       /* call out of order (e.g. step before set_potential) */

#define TDSE_FP32 0 /* complex64 state, FP64 phase arguments and reductions */
#define TDSE_FP64 1 /* complex128 state */

/* Plan flags.
 * TDSE_FLAG_UNFUSED_BASELINE: FP64, 2D/3D only.  Runs the reference CPU data
 * flow one GPU pass per operation (full complex128 half-kick and kinetic
 * tables, per-axis forward FFT sweeps, conjugation-identity inverse with
 * separate conj passes, separate mask pass; core/src/propagator.ixx:110-119,
 * core/src/fft.ixx:102-188).  It is the unfused comparison point of the
 * fused path, not a production mode.
 * TDSE_FLAG_IMAGINARY_TIME: the plan relaxes instead of propagating
 * (ImaginaryTimePropagator1D/3D, core/src/imaginary_time.ixx:21-144): dt is
 * dtau, the weights are real (exp(-V dtau/2), exp(-k^2 dtau/2m)) and every
 * tdse_step step ends with psi /= sqrt(sum |psi|^2 dV) (left unchanged for a
 * zero field, core/src/field.ixx:78-87).  No mask, no batching. */
#define TDSE_FLAG_UNFUSED_BASELINE 1u
#define TDSE_FLAG_IMAGINARY_TIME 2u
/* TDSE_FLAG_LATTICE: the 2D finite-difference propagator with Peierls link
 * phases (PeierlsLattice2D, core/src/lattice2d.ixx:31-275) on an n x m x 1 grid
 * (any n <= 2048, m; open boundaries, hopping 1/(2 h^2) per axis, onsite
 * V + 2 tx + 2 ty).  tdse_step = PeierlsLattice2D::step; with
 * TDSE_FLAG_IMAGINARY_TIME, tdse_step = relax (renormalised each step).  The
 * energy observable is the lattice <H>; no momentum, drive, field of
 * tdse_plan_set_magnetic, batching or kernel variants. */
#define TDSE_FLAG_LATTICE 4u

typedef struct tdse_ctx tdse_ctx;
typedef struct tdse_plan tdse_plan;

/* Library version string and human-readable status text. */
const char* tdse_version(void);
const char* tdse_status_string(int status);
/* Text of the last error reported on the calling thread ("" if none). */
const char* tdse_last_error(void);

/* Device context.  hbm_budget_bytes caps the device memory all plans of this
 * context may hold (0 = no cap); plan creation fails with TDSE_E_OOM beyond it
 * (the admission policy of core/src/vram_budget.ixx, applied to bytes). */
int tdse_ctx_create(int device_ordinal, uint64_t hbm_budget_bytes, tdse_ctx** out);
int tdse_ctx_destroy(tdse_ctx* ctx);
/* Bytes currently held by the context's plans. */
int tdse_ctx_bytes_in_use(const tdse_ctx* ctx, uint64_t* out);
/* Device memory a plan of this shape and precision would hold (for admission). */
int tdse_plan_bytes(const int n[3], int precision, int with_mask, uint64_t* out);

/* A propagator for one grid: replaces the SplitOperator1D/3D constructor
 * (core/src/propagator.ixx:23, 72).  lo/hi are the periodic box bounds per
 * axis (Grid1D::xmin/xmax), mass the effective mass (default 1.0 in the
 * reference), flags 0, TDSE_FLAG_UNFUSED_BASELINE, TDSE_FLAG_IMAGINARY_TIME and/or
 * TDSE_FLAG_LATTICE. */
int tdse_plan_create(tdse_ctx* ctx, const int n[3], const double lo[3], const double hi[3],
                     double dt, double mass, int precision, unsigned flags, tdse_plan** out);
int tdse_plan_destroy(tdse_plan* plan);

/* The potential V (n[0]*n[1]*n[2] doubles, x-fastest).  Builds the half-kick
 * phase -0.5*V*dt per cell (core/src/propagator.ixx:28-31, 78-81), reduced
 * modulo 2*pi in FP64 (then rounded to float on the FP32 path). */
int tdse_plan_set_potential(tdse_plan* plan, const double* v);
/* Real absorbing mask multiplied into psi after every step (null clears it).
 * core/src/potential.ixx:194-233 builds such masks: the product of per-axis
 * ramps.  A 3D or 2D mask equal bit for bit to such a product is applied by
 * the x sweeps from its ramps (no per-cell mask read per step); any other
 * mask is read per cell (4 or 8 more bytes per cell and step). */
int tdse_plan_set_mask(tdse_plan* plan, const double* mask);

/* State transfer, interleaved complex128 on the host. */
int tdse_state_upload(tdse_plan* plan, const double* psi);
int tdse_state_download(tdse_plan* plan, double* psi);
/* The state as interleaved complex64 (two floats per point, x-fastest): the
 * FP32 state as stored, the FP64 state rounded to nearest.  Waits for the
 * plan's queued work.  The reference engine's readback(std::vector<float>&),
 * solver/src/vk_engine.ixx:2136-2155 (half the bytes of tdse_state_download). */
int tdse_state_download_f32(tdse_plan* plan, float* out);
/* Copy the state of `src` into `dst` (same shape and precision), on the GPU:
 * by the runtime's device copy, or by a plain streaming copy kernel (16 bytes
 * per lane), the bandwidth reference the sweeps are measured against.  Both
 * queue on dst's stream after src's queued work. */
int tdse_state_copy(tdse_plan* dst, const tdse_plan* src);
int tdse_state_copy_kernel(tdse_plan* dst, const tdse_plan* src);
/* Device address of the resident state (complex64 or complex128 per point). */
int tdse_state_device_ptr(tdse_plan* plan, void** out);

/* nsteps Strang steps (core/src/propagator.ixx:44-53, 110-119), each followed
 * by the mask when one is set.  nsteps == 0 is the identity.  Asynchronous. */
int tdse_step(tdse_plan* plan, int nsteps);
/* Wait for the plan's queued work; timeout_ms < 0 waits forever. */
int tdse_plan_sync(tdse_plan* plan, int timeout_ms);
/* The plan's HIP stream (hipStream_t) for event timing by the caller. */
int tdse_plan_stream(tdse_plan* plan, void** out);

/* Observables of the resident state, FP64 results, fixed-order reductions.
 * One read pass gives norm, position, width and <V> (every call); the k-space
 * terms take one read-only line-FFT sweep per axis (no copy of the state).
 *   TDSE_OBS_NORM      norm = sum |psi|^2 * cell volume      (core/src/field.ixx:70-76)
 *   TDSE_OBS_POSITION  <x>, <y>, <z> (ratio form) and the     (core/src/observables.ixx:134-175)
 *                      widths sqrt(max(0, <x^2> - <x>^2))
 *   TDSE_OBS_ENERGY    <V> + <k^2/2m>, each term normalised  (core/src/observables.ixx:202-236)
 *   TDSE_OBS_MOMENTUM  <k_x>, <k_y>, <k_z> (ratio form)      (core/src/observables.ixx:52-65, 177-200)
 *   TDSE_OBS_VARIANCE  Var(H) = ||(H - <H>) psi||^2 / ||psi||^2, residual form, 1D grids
 *                      only (TDSE_E_UNSUPPORTED otherwise)    (core/src/observables.ixx:92-119)
 *   TDSE_OBS_PEAK      peak = max |psi|^2 over the cells       (solver/src/vk_engine.ixx:1665-1691,
 *                      (no cell volume)                        Engine::norm_and_peak)
 */
#define TDSE_OBS_NORM 1u
#define TDSE_OBS_POSITION 2u
#define TDSE_OBS_ENERGY 4u
#define TDSE_OBS_MOMENTUM 8u
#define TDSE_OBS_VARIANCE 16u
#define TDSE_OBS_PEAK 32u
typedef struct tdse_observables {
    double norm;
    double sum_abs2;      /* sum |psi|^2 without the cell volume */
    double position[3];
    double potential;     /* <V> */
    double kinetic;       /* <k^2 / 2m> */
    double energy;        /* potential + kinetic */
    double sigma[3];      /* sqrt(max(0, <x_a^2> - <x_a>^2)) */
    double momentum[3];   /* <k_a> */
    double energy_variance; /* TDSE_OBS_VARIANCE */
    double peak;          /* TDSE_OBS_PEAK */
} tdse_observables;
int tdse_observe(tdse_plan* plan, unsigned what, tdse_observables* out);
/* The position-space part of tdse_observe -- norm, sum |psi|^2, <x>, <y>, <z>,
 * the widths and <V> (kinetic, energy and momentum are set to 0) -- over a
 * state the caller holds in device memory (psi: complex64 or complex128 per
 * point by `precision`, x-fastest; v: the potential in the same precision, or
 * null), written to `out` in DEVICE memory, queued on `stream` (a
 * hipStream_t; null = the null stream), with no synchronisation.  No
 * counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, serving loops, harnesses that forbid
 * hipMalloc).  It allocates nothing: the reduction's partials live in a
 * static 147,528-byte device array per GPU, held until the process ends.
 * Calls that overlap on one GPU race on that array (order them on one
 * stream).  States of 4 GiB or more return TDSE_E_UNSUPPORTED.  Same block
 * count and summation order as tdse_observe on a plan of that shape. */
int tdse_observe_noalloc(const void* psi, const void* v, int precision, const int n[3],
                         const double lo[3], const double hi[3], tdse_observables* out,
                         void* stream);
/* Probability lo <= coord_axis < hi times the cell volume, not normalised
 * (core/src/observables.ixx:121-131, extended to one axis of a 3D grid). */
int tdse_probability_in_slab(tdse_plan* plan, int axis, double lo, double hi, double* out);
/* Probability lo[a] <= coord_a < hi[a] on every axis of extent above 1 (axes
 * of extent 1 are not restricted; infinite bounds leave an axis open) times
 * the cell volume, not normalised: a box region, the 3D form of
 * core/src/observables.ixx:121-131.  One restricted axis gives exactly
 * tdse_probability_in_slab. */
int tdse_probability_in_box(tdse_plan* plan, const double lo[3], const double hi[3], double* out);
/* Multiply the state by s on the GPU (renormalisation; solver/shaders/scale.comp.slang). */
int tdse_state_scale(tdse_plan* plan, double s);

/* Per-launch timing of `nsteps` steps: every sweep kernel on the plan's
 * stream is timed by HIP events recorded at its own start and end (the norm
 * finish of a relax step by events around its launch).  No reference
 * counterpart (the reference times whole steps, bench/bench_main.cpp:21-29).
 * ms_per_sweep[i] / launches[i] describe sweep
 * slot i of the step schedule (slots listed by tdse_plan_sweep_name); at most
 * max_slots are written; *n_slots receives the number of slots. */
int tdse_profile_step(tdse_plan* plan, int nsteps, int max_slots, double* ms_per_sweep,
                      int* launches, int* n_slots);
const char* tdse_plan_sweep_name(const tdse_plan* plan, int slot);
/* Algorithmic HBM bytes one steady-state step moves (state in and out per
 * sweep, kick phase and mask once). */
int tdse_plan_step_bytes(const tdse_plan* plan, uint64_t* out);

/* Dipole drive V(r, t) = amplitude (axis . r) cos(omega t) (core/src/drive.ixx:
 * 19-66): with a drive set, tdse_step(plan, n) is driven_step(psi, prop, d,
 * t0, n) with t0 the plan's clock.  The drive's half-kicks are diagonal like
 * the potential's, so they merge into the same kick sweeps (at no extra
 * memory traffic); amplitude 0 removes the drive (the static step, bit for
 * bit).  The clock starts at 0, advances by dt per step of every tdse_step
 * call and can be set.  Driven plans do not batch.  A drive whose kick angle
 * |amplitude| dt |axis . r| reaches 1e5 rad anywhere in the box returns
 * TDSE_E_UNSUPPORTED (the kernels reduce it with a short Cody-Waite step). */
int tdse_plan_set_drive(tdse_plan* plan, const double axis[3], double amplitude, double omega);
int tdse_plan_set_time(tdse_plan* plan, double t);
/* Uniform magnetic field `bfield` along `axis` (0 = x, 1 = y, 2 = z) in the
 * symmetric gauge (MagneticPropagator3D, core/src/magnetic.ixx:19-80): the
 * diamagnetic B^2/8 |r_perp|^2 joins the potential, and each step is
 * R(a) . Strang step . R(a), a = (B/2)(dt/2), with R the exact three-shear
 * rotation (core/src/rotation.ixx:73-84), one line-FFT sweep per shear; a
 * mask follows each whole step.  Call before tdse_plan_set_potential (or
 * tdse_plan_build); 3D grids, or 2D grids with axis 2.  0 removes the field.
 * No drive, imaginary time or batching with a field. */
int tdse_plan_set_magnetic(tdse_plan* plan, double bfield, int axis);

/* Lattice links (TDSE_FLAG_LATTICE plans).  A uniform field B along z in the
 * Landau gauge: x links of row j carry exp(-i B hx y_j)
 * (core/src/lattice2d.ixx:93-105).  A solenoid of flux phi in the cell
 * holding (xs, ys), string gauge: the x links of that column above it carry
 * exp(-i phi) (cut_up) or those below and at it exp(+i phi)
 * (core/src/lattice2d.ixx:60-91).  Each call replaces the links.  A uniform
 * field whose link angles |B hx y| reach 1e5 rad returns TDSE_E_UNSUPPORTED. */
int tdse_lattice_set_uniform_field(tdse_plan* plan, double b);
int tdse_lattice_set_solenoid(tdse_plan* plan, double phi, double xs, double ys, int cut_up);
int tdse_plan_get_time(const tdse_plan* plan, double* out);

/* Imaginary time (plans created with TDSE_FLAG_IMAGINARY_TIME).
 * tdse_step(plan, n) = ImaginaryTimePropagator*::relax(psi, n)
 * (core/src/imaginary_time.ixx:43-53, 112-122).
 * tdse_relax_deflated = relax_deflated(psi, lower, n)
 * (core/src/imaginary_time.ixx:124-141): after each step, psi -= <phi|psi> phi
 * for every lower state phi in order (<a|b> = sum conj(a) b dV,
 * core/src/field.ixx:88-99), then normalise.  Lower states are plans of the
 * same shape and precision (their current states).
 * tdse_plan_relax_stats: the norm before renormalisation of the last step, the
 * energy estimate -ln(norm_sq) / (2 dtau) (the norm of a normalised state
 * decays as exp(-2 E dtau), solver/src/vk_engine.ixx:3017-3023) and the peak
 * density max |psi|^2 / norm_sq of that step (after a deflated call: of the
 * step before projection). */
typedef struct tdse_relax_stats {
    double norm_sq;
    double energy;
    double peak;
} tdse_relax_stats;
int tdse_relax_deflated(tdse_plan* plan, const tdse_plan* const* lower, int nlower, int nsteps);
int tdse_plan_relax_stats(tdse_plan* plan, tdse_relax_stats* out);

/* Kernel variants.  Each sweep of the step schedule (slot, as above) can run
 * one of up to 4 compiled variants of its kernel: the base or an alternative
 * radix list (a different number of points per thread) times one or two
 * times the lines per block.  Variant 0 is the default of every slot.
 * tdse_plan_autotune times the compiled variants on a scratch copy of the
 * state (allocated on first use, counted in the context budget; the state
 * itself is left unchanged) and keeps the fastest: the sweeps of every step
 * by the time of whole steps (at least `reps` steps and about 2 ms per
 * measurement, one pass of coordinate descent over the slots), the sweeps run
 * once per call by the time of `reps` launches of their own.  A change from
 * the current choice must be at least 2% faster.  variants[] receives the
 * choice per slot.  Lengths 64-1024, 96 and 192 have variants; the others and the
 * one-block 1D kernel run variant 0.  A batch sweeps each member with the
 * variants its plan holds when the batch is created (plans of one line
 * length, mode and variant share a launch).  No reference counterpart (the
 * reference has one FFT plan per length, core/src/fft.ixx:64-100). */
int tdse_plan_autotune(tdse_plan* plan, int reps, int max_slots, int* variants, int* n_slots);
int tdse_plan_get_variants(const tdse_plan* plan, int max_slots, int* variants, int* n_slots);
/* TDSE_E_UNSUPPORTED when a slot's variant is not compiled for its sweep. */
int tdse_plan_set_variants(tdse_plan* plan, const int* variants, int n_slots);

/* Batches: many plans (any shapes, both precisions) advanced together.
 * Members of one class (precision, geometry, line lengths) form units whose
 * one-step working set stays within 64 MB (a quarter of the Infinity Cache);
 * a unit's members share one launch per sweep (per kernel variant), and each
 * of four streams runs its units' steps back to back, so a unit's state stays
 * cached between its sweeps while four units overlap on the device.
 * use_graph must be 0: replay of stream-captured hipGraphs is switched off
 * (TDSE_E_UNSUPPORTED).  The batch snapshots each plan's buffers at
 * creation: re-create it after tdse_plan_set_mask changes a plan's mask from
 * null to non-null.  tdse_batch_step(b, n) == tdse_step(p, n) for every plan
 * p of the batch (no reference counterpart: the reference steps one scene at a time).
 * Each member is swept in the kernel variants its plan holds at creation
 * (tdse_plan_autotune / tdse_plan_set_variants). */
typedef struct tdse_batch tdse_batch;
int tdse_batch_create(tdse_ctx* ctx, tdse_plan* const* plans, int count, tdse_batch** out);
int tdse_batch_destroy(tdse_batch* batch);
int tdse_batch_step(tdse_batch* batch, int nsteps, int use_graph);
int tdse_batch_sync(tdse_batch* batch, int timeout_ms);
int tdse_batch_stream(tdse_batch* batch, void** out);
/* Number of batched launches in one interior round of every unit. */
int tdse_batch_launches_per_round(const tdse_batch* batch, int* out);

/* Uneven submissions: advance plans[i] by nsteps[i] steps (tdse_step), each
 * plan on its own stream, so solves of different shapes, precisions and step
 * counts overlap on the GPU (the policy the served stream measured fastest).
 * Every plan must belong to ctx and have a potential; all arguments are
 * checked before anything is queued (a device error while queuing leaves the
 * earlier plans' steps queued).  Asynchronous.  No reference counterpart:
 * the reference GPU engine queues one scene's steps at a time
 * (solver/src/vk_engine.ixx:645, Engine::step_async). */
int tdse_batch_submit(tdse_ctx* ctx, tdse_plan* const* plans, const int* nsteps, int count);
/* Wait for the queued work of every live plan and batch of ctx;
 * timeout_ms < 0 waits forever, otherwise TDSE_E_TIMEOUT once timeout_ms
 * have passed with work still queued (0: a non-blocking poll).  Plans and
 * batches of ctx are not created or destroyed on other threads while it
 * waits (those calls wait for it).  Behavioural precedent: the engine's
 * bounded wait, solver/src/vk_engine.ixx:681 (Engine::wait_async). */
int tdse_sync(tdse_ctx* ctx, int timeout_ms);

/* Build a plan's inputs on the GPU from a recipe: potential (and its kick
 * phase), optional absorbing mask, and the normalised Gaussian packet as the
 * state.  Same formulas as the host builders below, evaluated in FP64 on the
 * device (equal to a few ulps, not bit for bit).  Replaces
 * set_potential + set_mask + state_upload without host arrays. */
#define TDSE_POT_NONE 0
#define TDSE_POT_HARMONIC 1            /* 0.5 strength^2 |r - center|^2 */
#define TDSE_POT_REGULARIZED_COULOMB 2 /* -strength / |r - center|, nucleus cell average */
#define TDSE_POT_COULOMB_FLOOR 3       /* -strength / max(|r - center|, h/2) */
#define TDSE_POT_BARRIER 4             /* strength on x_lo <= x < x_hi */
#define TDSE_POT_SOFT_COULOMB 5        /* -strength / sqrt(|r - center|^2 + softening^2) */
typedef struct tdse_recipe {
    int potential;
    double strength;
    double center[3];
    double x_lo, x_hi;
    double mask_width;                 /* <= 0: no mask */
    double r0[3], sigma[3], k0[3];     /* Gaussian packet (sigma: density std-dev) */
    double softening;                  /* TDSE_POT_SOFT_COULOMB */
} tdse_recipe;
int tdse_plan_build(tdse_plan* plan, const tdse_recipe* recipe);

/* Host input builders, FP64, x-fastest, bit-identical to the reference
 * builders for the same arguments (n x 1 x 1 grids use the 1D forms).
 *   harmonic            0.5 w^2 |r - c|^2           core/src/potential.ixx:16-48
 *   barrier             v0 on x_lo <= x < x_hi      core/src/potential.ixx:102-132
 *   regularized_coulomb -Z/r, nucleus cell -2.3800774 Z/h   core/src/potential.ixx:134-154
 *   coulomb_floor       -Z / max(r, h/2)            (no reference builder)
 *   soft_coulomb        -Z / sqrt(r^2 + a^2)        core/src/potential.ixx:25-33, 51-65
 *   absorbing_mask      separable sin^2 ramp        core/src/potential.ixx:194-233
 *   gaussian            normalised packet, interleaved complex128 out
 *                                                   core/src/wavepacket.ixx:16-48 */
int tdse_build_harmonic(const int n[3], const double lo[3], const double hi[3], double omega,
                        const double c[3], double* v);
int tdse_build_barrier(const int n[3], const double lo[3], const double hi[3], double v0,
                       double x_lo, double x_hi, double* v);
int tdse_build_regularized_coulomb(const int n[3], const double lo[3], const double hi[3],
                                   double z, const double c[3], double* v);
int tdse_build_coulomb_floor(const int n[3], const double lo[3], const double hi[3], double z,
                             const double c[3], double* v);
int tdse_build_soft_coulomb(const int n[3], const double lo[3], const double hi[3], double z,
                            double a, const double c[3], double* v);
int tdse_build_absorbing_mask(const int n[3], const double lo[3], const double hi[3],
                              double width, double* m);
int tdse_build_gaussian(const int n[3], const double lo[3], const double hi[3], const double r0[3],
                        const double sigma[3], const double k0[3], double* psi);
/* The reference propagator's FP64 phase tables, interleaved complex128,
 * x-fastest, bit-identical to SplitOperator1D/3D::half_potential_phase() and
 * kinetic_phase() (core/src/propagator.ixx:23-40, 72-101, 107-108):
 * half_v = cis(-0.5 V dt) per cell; kinetic = cis(-0.5 k^2 dt) on an n x 1 x 1
 * grid, cis(-0.5 (kx^2 + ky^2 + kz^2) / mass dt) over the full grid otherwise.
 * Either output may be null (v only needed for half_v).  The GPU path does
 * not use these full tables (it keeps per-axis kinetic tables and reduced
 * kick angles); they are for callers that read the reference's tables
 * (core/src/drive.ixx:58-62, tests/propagator_tables_test.cpp:28-65). */
int tdse_build_phase_tables(const int n[3], const double lo[3], const double hi[3], const double* v,
                            double dt, double mass, double* half_v, double* kinetic);

#ifdef __cplusplus
}
#endif

#endif /* TDSE_H */
