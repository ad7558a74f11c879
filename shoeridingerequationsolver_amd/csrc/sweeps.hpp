// sweeps.hpp -- the fused line-sweep kernels of the split-operator step.
//
// One Strang step  psi <- e^{-iV dt/2} IFFT e^{-ik^2 dt/2m} FFT e^{-iV dt/2} psi
// (core/src/propagator.ixx:44-53, 110-119) is run as a few full-array sweeps,
// each owning whole lines of one axis:
//
//   3D:  X[ IFFT_x . mask . kick . FFT_x ]  Y[ FFT_y ]  Z[ FFT_z . K . IFFT_z ]  Y[ IFFT_y ]
//   2D:  X[ IFFT_x . mask . kick . FFT_x ]  Y[ FFT_y . K . IFFT_y ]
//   1D, N = N1*N2 (four-step):
//        COL[ IFFT_n2 . mask . kick . FFT_n2 . twiddle ]  ROW[ FFT_n1 . K' . IFFT_n1 . twiddle* ]
//   1D, N <= 1024: one block per line runs every step in registers.
//
// The interior kicks of consecutive steps merge (diagonal factors commute):
// mask * e^{-iV dt}; the first sweep of a call applies e^{-iV dt/2} only and
// the last one e^{-iV dt/2} followed by the mask (the caller-side per-step
// mask of scenario/src/line1d_director.ixx:202-211).  The inverse's 1/N is
// folded into K.  Between calls psi is kept in position space, natural order.
//
// Phases are applied in FP64 for both state precisions: the kick is
// m * exp(i f phi) with phi = -V dt / 2 stored per cell, reduced modulo 2 pi
// in FP64 on the host (and rounded to float on the FP32 path), and the kinetic factor is
// the product of per-axis FP64 tables exp(-i k_a^2 dt / 2m) (the reference
// builds the same phase from the summed angle, core/src/propagator.ixx:83-98;
// the two agree to rounding).  The state point is widened to FP64, multiplied
// and rounded once, so a phase never enters with a float-rounded magnitude.
//
// Global memory is addressed through buffer resources: one SGPR descriptor
// per array and 32-bit byte offsets split into a per-lane part (VGPR) and a
// per-radix-slot part that is uniform across the wave (SGPR).  Arrays are
// therefore limited to 4 GiB each.  Reads beyond the descriptor's range
// return zero and writes beyond it are dropped by the hardware.
#pragma once

#include <stdint.h>

#include "line_plans.hpp"

// Minimum waves per SIMD the compiler must fit (register cap); 1 = no cap.
#ifndef TDSE_MIN_WAVES
#define TDSE_MIN_WAVES 1
#endif

namespace tdse {

enum SweepMode : int {
    // contiguous (unit-stride lines)
    C_KICK_FIRST = 0,  // kick_half, FFT
    C_KICK_MID = 1,    // IFFT, mask*kick_full, FFT
    C_KICK_LAST = 2,   // IFFT, kick_half*mask
    C_KIN_TW = 3,      // FFT, K (four-step order), IFFT, conj twiddle
    C_SINGLE = 4,      // every step of a 1D line in one block
    // strided
    S_FWD = 5,           // FFT
    S_INV = 6,           // IFFT
    S_KIN = 7,           // FFT, K, IFFT
    S_KICK_FIRST_TW = 8, // kick_half, FFT, twiddle
    S_KICK_MID_TW = 9,   // IFFT, mask*kick_full, FFT, twiddle
    S_KICK_LAST = 10,    // IFFT, kick_half*mask
    C_FWD = 11,          // FFT only (observables: psi -> k space)
    S_FWD_TW = 12,       // FFT, twiddle (observables, four-step 1D)
    // imaginary time: real weights w = exp(-V dtau / 2) and a norm per step
    C_RELAX_FIRST = 13,     // w, FFT
    C_RELAX_MID = 14,       // IFFT, scale * w, measure, w, FFT
    C_RELAX_LAST = 15,      // IFFT, scale * w, measure
    C_RELAX_SINGLE = 16,    // every relax step of a 1D line in one block
    S_RELAX_FIRST_TW = 17,  // w, FFT, twiddle
    S_RELAX_MID_TW = 18,    // IFFT, scale * w, measure, w, FFT, twiddle
    S_RELAX_LAST = 19,      // IFFT, scale * w, measure
    // observables: read-only line spectra (no store)
    C_OBS = 20,             // FFT, accumulate |psi_k|^2 (1, k, k^2)
    S_OBS = 21,             // FFT, accumulate |psi_k|^2 (1, k, k^2)
    // shears of a three-shear rotation: FFT, exp(-i k d(line)) / N, IFFT (, mask)
    C_SHEAR = 22,
    S_SHEAR = 23,
    // the magnetic step's first and last x sweeps fused with the adjacent x
    // shear of the rotation around it
    C_SHEAR_KICK_FIRST = 24,  // FFT, shear, IFFT, kick_half, FFT
    C_KICK_LAST_SHEAR = 25,   // IFFT, kick_half, FFT, shear, IFFT (, mask)
    kNumModes = 26
};

// A kick mode with the dipole drive (driven plans): the mode | kDrive.  The
// drive's extra exponential lives only in these instantiations, so the
// static kick sweeps keep their register budget.
constexpr int kDrive = 32;
constexpr int base_mode(int m) { return m & ~kDrive; }
constexpr bool mode_driven(int m) { return (m & kDrive) != 0; }
constexpr bool mode_kicks(int m) {
    return m == C_KICK_FIRST || m == C_KICK_MID || m == C_KICK_LAST || m == C_SINGLE ||
           m == S_KICK_FIRST_TW || m == S_KICK_MID_TW || m == S_KICK_LAST;
}

constexpr bool mode_contiguous(int mode) {
    const int m = base_mode(mode);
    return m <= C_SINGLE || m == C_FWD || (m >= C_RELAX_FIRST && m <= C_RELAX_SINGLE) ||
           m == C_OBS || m == C_SHEAR || m == C_SHEAR_KICK_FIRST || m == C_KICK_LAST_SHEAR;
}
constexpr bool mode_observes(int m) { return m == C_OBS || m == S_OBS; }
// modes that measure the norm across blocks (deferred renormalisation)
constexpr bool mode_measures(int m) {
    return m == C_RELAX_MID || m == C_RELAX_LAST || m == S_RELAX_MID_TW || m == S_RELAX_LAST;
}

// Device state of an imaginary-time plan.  A relax step ends with
// psi /= sqrt(norm); the norm is only known once the sweep that measures it
// has finished, so the division is deferred: the next measuring sweep
// multiplies by the scale the previous one left (by step parity), and the
// call ends with one scaling pass.
struct RelaxSlots {
    double scale[2];      // 1 / sqrt(norm) of the previous measured step, by parity
    double norm_sq;       // sum |psi|^2 dV before renormalisation, last measured step
    double peak;          // max |psi|^2 there (before renormalisation)
    unsigned int nblocks;  // blocks (partials) of the last measuring sweep
    unsigned int pad;
};

// Blocks of a measuring sweep (partials hold two doubles per block): one
// tile per block at 256^3 in FP32 (measured 1% faster than two), two in FP64
// (measured 0.4% faster than one).
constexpr int kRelaxMaxBlocks = 4096;
template <typename T>
constexpr int relax_max_blocks() { return sizeof(T) == 4 ? kRelaxMaxBlocks : kRelaxMaxBlocks / 2; }

// Per-thread accumulation over all tiles of a block: a measuring sweep keeps
// (sum |psi|^2, max |psi|^2), an observing sweep (sum w, sum k w, sum k^2 w)
// of its line spectra.
struct SweepAcc {
    double q0 = 0.0;
    double q1 = 0.0;
    double q2 = 0.0;
};

// Blocks of an observing sweep (partials hold three doubles per block).
constexpr int kObsMaxBlocks = 2048;

struct SweepArgs {
    void* psi;            // cx<T>*: the state, x-fastest
    const void* phi;      // T*: half-kick angle -V dt/2 per cell
    const void* mask;     // T*: per-step absorbing mask, or null
    const void* tw;       // cx<double>*: W_N^m = exp(-2 pi i m / N), m < N, this line length
    uint32_t psi_bytes;   // byte extent of psi (buffer range)
    uint32_t real_bytes;  // byte extent of phi / mask
    long long nlines;     // contiguous: number of lines
    uint32_t wstride;     // strided: element stride of the outer index w
    uint32_t estride;     // strided: element stride along the line
    int nx;               // strided: extent of the unit-stride index x
    int nw;               // strided: extent of the outer index w
    // kinetic factor: product of per-axis FP64 tables exp(-i k^2 dt / 2m),
    // indexed by slot (x, w, p); a null slot contributes 1
    const void* kt[3];
    double scale;         // 1 / total points (inverse normalisation)
    int n_total;          // four-step twiddle modulus (1D), else 0
    int n_rows;           // four-step: number of rows N2 (contiguous K mapping)
    int nsteps;           // C_SINGLE / C_RELAX_SINGLE: steps to run
    // imaginary time
    RelaxSlots* relax;    // device slots, or null
    double* partials;     // per-block (sum, max) of a measuring sweep
    double cell;          // volume element dV
    int parity;           // measuring sweeps: read scale[parity], write scale[parity ^ 1]
    double dk;            // observing sweeps: 2 pi / L of the line axis
    // dipole drive, a time-dependent linear kick: angle += dg * (axis . r)
    double dg;            // minus the drive angles theta merged into this kick (0: none)
    // axis . r of grid point (i, j, k) = du_c + du_x i + du_y j + du_z k
    double du_c, du_x, du_y, du_z;
    int dn1;              // contiguous: lines per z index (ny); four-step: row length n1
    // C_SINGLE: theta(t) = damp cos(domega t) ddt / 2 with t = dt0 + s ddt
    double damp, domega, dt0, ddt;
    // shear sweeps: the line's shift d = sh_d0 + sh_d1 * idx, idx its index
    // along the coordinate axis (sh_sel 0: y of an x line / x of a strided
    // line; 1: z of an x line / the outer index w of a strided line)
    double sh_d0, sh_d1;
    int sh_sel;
    // absorbing mask as the product of per-axis ramps, (mx * my) * mz, the
    // way absorbing_mask builds it (core/src/potential.ixx:217-233): with
    // mask_sep, `mask` is the ramp along the line (stored form, one entry per
    // point of the line) and mask_yz the ramps of the line's other two axes
    // (dn1 entries, then the rest); the x sweep then reads no mask array
    const double* mask_yz;
    uint32_t mask_bytes;  // byte extent of mask
    int mask_sep;
    // the previous measuring sweep's norm, finished by block 0 of this sweep
    // before its tiles (fin_slots null: nothing to finish): sum and max of
    // fin_slots->nblocks partials -> scale[fin_parity ^ 1], norm_sq, peak
    const double* fin_partials;
    RelaxSlots* fin_slots;
    double fin_cell;
    int fin_parity;
};

// ---- buffer access -----------------------------------------------------------

constexpr int32_t kBufferConfig = 0x00020000;  // raw 32-bit data format, gfx9 family

template <typename V>
struct Buf {
    __amdgpu_buffer_rsrc_t r;
    __device__ __forceinline__ Buf(const void* p, uint32_t bytes)
        : r(__builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), static_cast<short>(0), bytes,
                                              kBufferConfig)) {}
    __device__ __forceinline__ V ld(uint32_t voff, uint32_t soff) const {
        if constexpr (sizeof(V) == 4) {
            return __builtin_bit_cast(V, __builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 0));
        } else if constexpr (sizeof(V) == 8) {
            return __builtin_bit_cast(V, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0));
        } else {
            return __builtin_bit_cast(V, __builtin_amdgcn_raw_buffer_load_b128(r, voff, soff, 0));
        }
    }
    __device__ __forceinline__ void st(V v, uint32_t voff, uint32_t soff) const {
        if constexpr (sizeof(V) == 4) {
            __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(uint32_t, v), r, voff, soff, 0);
        } else if constexpr (sizeof(V) == 8) {
            using U = __attribute__((ext_vector_type(2))) uint32_t;
            __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(U, v), r, voff, soff, 0);
        } else {
            using U = __attribute__((ext_vector_type(4))) uint32_t;
            __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(U, v), r, voff, soff, 0);
        }
    }
};

// ---- phase helpers ------------------------------------------------------------

// sin and cos of x for |x| <= 8 (kick angles are stored reduced to
// [-pi, pi] and doubled at most).  Cody-Waite reduction by pi/2 with a
// two-part constant (A = pi/2 rounded to 33 bits so q*A is exact, B = the
// rest rounded to double; the split is 3.5e-27 short of pi/2), then the
// Taylor series of sin and cos on |r| <= pi/4 to degree 17 / 18 (truncation
// below 1e-19), Horner in FP64.  About 30 FP64 operations, no table, no branch.
__device__ __forceinline__ void sincos_small(double x, double* s, double* c) {
    constexpr double kTwoOverPi = 0.6366197723675814;
    constexpr double kPio2A = 1.5707963267341256;   // 0x1.921fb544p+0
    constexpr double kPio2B = 6.077100506506192e-11;
    const double q = rint(x * kTwoOverPi);
    const double r = (x - q * kPio2A) - q * kPio2B;
    const double r2 = r * r;
    // sin r = r + r^3 P(r^2), P = sum_{k>=1} (-1)^k r^{2k-2} / (2k+1)!
    double p = 1.0 / 355687428096000.0;      //  1/17!
    p = fma(p, r2, -1.0 / 1307674368000.0);  // -1/15!
    p = fma(p, r2, 1.0 / 6227020800.0);      //  1/13!
    p = fma(p, r2, -1.0 / 39916800.0);       // -1/11!
    p = fma(p, r2, 1.0 / 362880.0);          //  1/9!
    p = fma(p, r2, -1.0 / 5040.0);           // -1/7!
    p = fma(p, r2, 1.0 / 120.0);             //  1/5!
    p = fma(p, r2, -1.0 / 6.0);              // -1/3!
    const double sr = fma(r * r2, p, r);
    // cos r = 1 + r^2 Q(r^2), Q = sum_{k>=1} (-1)^k r^{2k-2} / (2k)!
    double qc = -1.0 / 6402373705728000.0;   // -1/18!
    qc = fma(qc, r2, 1.0 / 20922789888000.0);  //  1/16!
    qc = fma(qc, r2, -1.0 / 87178291200.0);    // -1/14!
    qc = fma(qc, r2, 1.0 / 479001600.0);       //  1/12!
    qc = fma(qc, r2, -1.0 / 3628800.0);        // -1/10!
    qc = fma(qc, r2, 1.0 / 40320.0);           //  1/8!
    qc = fma(qc, r2, -1.0 / 720.0);            // -1/6!
    qc = fma(qc, r2, 1.0 / 24.0);              //  1/4!
    qc = fma(qc, r2, -0.5);                    // -1/2!
    const double cr = fma(r2, qc, 1.0);
    // x = q pi/2 + r: sin = (sr, cr, -sr, -cr), cos = (cr, -sr, -cr, sr) for q mod 4
    const int qi = static_cast<int>(q) & 3;
    const double ss = (qi & 1) ? cr : sr;
    const double cc = (qi & 1) ? sr : cr;
    *s = (qi & 2) ? -ss : ss;
    *c = ((qi + 1) & 2) ? -cc : cc;
}

// Absorbing-mask storage: FP64 plans store m, FP32 plans store 1 - m, so a
// value just below 1 (the inner edge of the absorbing layer, where most of
// the absorption happens) keeps its full precision as a small complement
// instead of rounding to 1.
template <typename T>
__device__ __forceinline__ double mask_value(T stored) {
    if constexpr (sizeof(T) == 8) return stored;
    else return 1.0 - static_cast<double>(stored);
}
template <typename T>
constexpr T kMaskOne = sizeof(T) == 8 ? T(1) : T(0);  // stored form of m = 1

// exp(i f phi) from t = tan(phi / 2), which FP32 plans store in place of
// phi: exp(i phi) = ((1 - t^2) + 2 i t) / (1 + t^2), of modulus 1 up to a few
// FP64 roundings and with no sine or cosine; f = 2 squares it.  The
// reciprocal is the hardware estimate r0 (about 17 bits) refined by one
// third-order step, r0 (1 + e + e^2) with e = 1 - den r0 (relative error
// about e^3, below 2^-50).  (One Newton step instead, error e^2 and always
// short, raised the 2000-step FP32 norm drift from 4e-8 to 2.3e-7.)
__device__ __forceinline__ cx<double> tan_rotation(double t, double f) {
    const double t2 = t * t;
    const double den = 1.0 + t2;
    double r = __builtin_amdgcn_rcp(den);
    const double e = fma(-den, r, 1.0);
    r = fma(r, fma(e, e, e), r);
    const double c = (1.0 - t2) * r, s = (t + t) * r;
    if (f == 2.0) return {fma(c, c, -s * s), (c + c) * s};
    return {c, s};
}

// kick factor m * exp(i * (f * phi + extra)); f = 1 (half kick) or 2 (merged
// full kick), extra = the drive's linear angle (0 without one).  Real-time
// plans of both precisions store tan(phi / 2) (tan_rotation).
template <typename T>
__device__ __forceinline__ cx<double> kick_factor(T stored, double f, double m, double extra,
                                                  bool drive) {
    cx<double> k = tan_rotation(static_cast<double>(stored), f);
    if (drive) {
        double s, c;
        sincos_small(extra, &s, &c);
        k = {k.x * c - k.y * s, k.x * s + k.y * c};
    }
    return {k.x * m, k.y * m};
}

// theta(t) = E0 cos(w t) dt / 2 of a dipole drive (core/src/drive.ixx:25-26)
__device__ __forceinline__ double drive_theta(const SweepArgs& a, double t) {
    return a.damp * cos(a.domega * t) * 0.5 * a.ddt;
}

// four-step twiddle exp(DIR * 2 pi i (a*b mod n) / n) in FP64
template <int DIR>
__device__ __forceinline__ cx<double> step_twiddle(long long a, long long b, int n) {
    const long long m = (a * b) % n;
    double s, c;
    sincospi(2.0 * static_cast<double>(m) / static_cast<double>(n), &s, &c);
    return {c, DIR * s};
}

// ---- LDS exchange buffers -------------------------------------------------------

// Contiguous tiles: every line lies inside one wave (its N / E threads are
// consecutive and N / E divides 64), so a line's exchange only has to order
// the wave's own LDS accesses, which execute in order: the barrier fences the
// compiler, not the block.  Waves of a block then run their tiles
// independently and drift apart, so one wave's loads and stores overlap
// another's arithmetic (x sweep at 256^3 FP32: 59.1 -> 58.0 us against block
// barriers, same box).
template <typename T, bool TW_TABLE = false>
struct ContigX {
    static constexpr bool kTwTable = TW_TABLE;
    cx<T>* base;  // this thread's line
    __device__ __forceinline__ void sync() const {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
    __device__ __forceinline__ void put(int i, cx<T> z) const { base[lds_pad(i)] = z; }
    __device__ __forceinline__ cx<T> get(int i) const { return base[lds_pad(i)]; }
};

// Strided tiles: a line's threads span the block's waves (the unit-stride
// index is fastest across lanes, for whole cache lines per access).
template <typename T, int TX>
struct StridedX {
    static constexpr bool kTwTable = false;
    cx<T>* base;  // lds + xi
    __device__ __forceinline__ void sync() const { __syncthreads(); }
    __device__ __forceinline__ void put(int i, cx<T> z) const { base[lds_pad(i) * TX] = z; }
    __device__ __forceinline__ cx<T> get(int i) const { return base[lds_pad(i) * TX]; }
};

template <int N>
__device__ __forceinline__ void stage_table(cx<double>* lds, const void* g, int tid, int nthreads) {
    const cx<double>* src = static_cast<const cx<double>*>(g);
    for (int i = tid; i < N; i += nthreads) lds[i] = src[i];
}

// ---- the per-line work of every mode ------------------------------------------------

// One line (or one lane's share of it) of a sweep.  Point (b, r) of the
// thread sits at element  base + (j_b + r*N/R) * es  of psi/phi/mask.
template <typename T, int N, class X, int V = 0>
struct Line {
    using S = Shape<T, N, V>;
    using Fwd = typename S::P::Fwd;
    using Inv = Reverse<Fwd>;
    static constexpr int E = S::E;
    static constexpr int R1 = First<Fwd>::value, RL = Last<Fwd>::value;
    static constexpr uint32_t CS = sizeof(cx<T>), RS = sizeof(T);

    const SweepArgs& a;
    const Buf<cx<T>>& P;
    const Buf<T>& F;
    const Buf<T>& M;
    X& x;
    const cx<double>* tw;
    int t;
    uint32_t base, es;
    bool valid;
    cx<T> v[E];

    template <int R>
    __device__ __forceinline__ void load() {
        if (valid) {
            for_each_point<N, E, R>(t, [&](int s, int, int j, int ro) {
                v[s] = P.ld((base + j * es) * CS, ro * es * CS);
            });
        } else {
            for (int s = 0; s < E; ++s) v[s] = cx<T>{0, 0};
        }
    }
    template <int R>
    __device__ __forceinline__ void store() {
        if (!valid) return;
        for_each_point<N, E, R>(t, [&](int s, int, int j, int ro) {
            P.st(v[s], (base + j * es) * CS, ro * es * CS);
        });
    }

    // Kick inputs in pattern R1 (position space after IFFT / before FFT):
    // all loads are issued together, so their latency is paid once per line.
    // (Issuing them before the inverse transform measured no faster.)
    T kp[E];  // kick phase (FP32 real-time plans: its half-angle tangent)
    T km[E];  // mask, stored form (see mask_value)
    __device__ __forceinline__ void load_kick(bool with_mask) {
        if (!valid) return;
        // the branches are uniform: each case issues its loads back to back
        if (with_mask && a.mask && a.mask_sep) {
            for_each_point<N, E, R1>(t, [&](int s, int, int j, int ro) {
                const uint32_t vo = (base + j * es) * RS, so = ro * es * RS;
                kp[s] = F.ld(vo, so);
                km[s] = M.ld(j * es * RS, so);
            });
        } else if (with_mask && a.mask) {
            for_each_point<N, E, R1>(t, [&](int s, int, int j, int ro) {
                const uint32_t vo = (base + j * es) * RS, so = ro * es * RS;
                kp[s] = F.ld(vo, so);
                km[s] = M.ld(vo, so);
            });
        } else {
            for_each_point<N, E, R1>(t, [&](int s, int, int j, int ro) {
                kp[s] = F.ld((base + j * es) * RS, ro * es * RS);
                km[s] = kMaskOne<T>;
            });
        }
    }
    // v *= m exp(i (f phi + g u)), the exponential in FP64; u = axis . r of
    // the point (du0 + du1 * position along the line), g the drive angle of
    // this kick (0: none).
    double du0 = 0.0, du1 = 0.0;
    // separable mask: the ramps of the line's other two axes (1 otherwise)
    double my = 1.0, mz = 1.0;
    // mask value of a point from its stored line-ramp entry: (m_x m_y) m_z,
    // the reference's product order, for FP64 states; m_x (m_y m_z) with the
    // line's factor formed once for FP32 states (an FP64 rounding apart, far
    // below the state's precision)
    __device__ __forceinline__ double mask_of(T stored) const {
        if constexpr (sizeof(T) == 8) return (mask_value(stored) * my) * mz;
        else return mask_value(stored) * (my * mz);
    }
    template <bool DRIVE>
    __device__ __forceinline__ void kick_points(double f, bool use_mask, double g) {
        if constexpr (DRIVE && sizeof(T) == 4) {
            // the static kick, then the drive's exp(i g u) along the line by
            // recurrence: the points of butterfly b sit at j_b + r N / R1
            // (r < R1), so exp(i g u) = exp(i g (du0 + du1 j_b))
            // exp(i g du1 N / R1)^r (FP64, at most 32 products: relative
            // error below 1e-14).  Two passes keep the registers of one.
            kick_points<false>(f, use_mask, 0.0);
            __builtin_amdgcn_sched_barrier(0);
            constexpr int TT = N / E;
            double ss, cs;
            sincos_small(g * du1 * static_cast<double>(N / R1), &ss, &cs);
#pragma unroll
            for (int b = 0; b < E / R1; ++b) {
                const int j = t + b * TT;
                double s0, c0;
                sincos_small(g * (du0 + du1 * j), &s0, &c0);
                cx<double> w = {c0, s0};
#pragma unroll
                for (int r = 0; r < R1; ++r) {
                    const int s = b * R1 + r;
                    v[s] = mul_d(v[s], w);
                    w = {w.x * cs - w.y * ss, w.x * ss + w.y * cs};
                }
            }
        } else {
            for_each_point<N, E, R1>(t, [&](int s, int, int j, int ro) {
                const double extra = DRIVE ? g * (du0 + du1 * (j + ro)) : 0.0;
                const double m = use_mask ? mask_of(km[s]) : 1.0;
                v[s] = mul_d(v[s], kick_factor<T>(kp[s], f, m, extra, DRIVE));
            });
        }
    }
    template <bool DRIVE>
    __device__ __forceinline__ void apply_kick(double f, bool use_mask, double g) {
        if (!valid) return;
        if constexpr (sizeof(T) == 4) {
            // FP32 states always take the exponential (its test for a wave
            // of zero phases cost an instruction per point and, in 3D,
            // never passes: every x line crosses the potential)
            __builtin_amdgcn_sched_barrier(0);
            kick_points<DRIVE>(f, use_mask, g);
        } else {
            // FP64 states: waves with no angle at all multiply by the real
            // m alone.  (Taking the exponential unconditionally here let the
            // compiler overlap it with the transform before it, at a wave
            // of occupancy in the x sweeps.)
            bool nz = DRIVE;
            for (int s = 0; s < E; ++s) nz |= kp[s] != T(0);
            if (__any(nz)) {
                kick_points<DRIVE>(f, use_mask, g);
            } else if (use_mask && a.mask) {
                for (int s = 0; s < E; ++s) v[s] = cscale(v[s], mask_of(km[s]));
            }
        }
    }
    template <bool DRIVE>
    __device__ __forceinline__ void kick(double f, bool with_mask, double g) {
        load_kick(with_mask);
        apply_kick<DRIVE>(f, with_mask, g);
    }
    // Imaginary time.  kp holds phi = -V dtau / 2 (not reduced); the weight
    // exp(phi) is formed in FP64.
    __device__ __forceinline__ void relax_first() {
        load_kick(false);
        if (!valid) return;
        for (int s = 0; s < E; ++s) {
            const double e = exp(static_cast<double>(kp[s]));
            v[s] = {static_cast<T>(e * v[s].x), static_cast<T>(e * v[s].y)};
        }
    }
    // v <- v * pre * w, accumulated into (sum, peak) of |v|^2, then (second)
    // times w again; one rounding per component.
    __device__ __forceinline__ void relax_measure(double pre, bool second, SweepAcc& ra) {
        load_kick(false);
        if (!valid) return;
        for (int s = 0; s < E; ++s) {
            const double e = exp(static_cast<double>(kp[s]));
            const double f = pre * e;
            double xr = f * v[s].x, xi = f * v[s].y;
            const double a2 = xr * xr + xi * xi;
            ra.q0 += a2;
            ra.q1 = a2 > ra.q1 ? a2 : ra.q1;
            if (second) {
                xr *= e;
                xi *= e;
            }
            v[s] = {static_cast<T>(xr), static_cast<T>(xi)};
        }
    }
    // Spectrum moments of the transformed line (pattern RL, bin i):
    // sum |v|^2, sum k |v|^2, sum k^2 |v|^2 with k the fftfreq wavenumber.
    __device__ __forceinline__ void observe_spectrum(SweepAcc& acc) {
        if (!valid) return;
        for_each_point<N, E, RL>(t, [&](int s, int i, int, int) {
            const double w = static_cast<double>(v[s].x) * v[s].x +
                             static_cast<double>(v[s].y) * v[s].y;
            const double k = a.dk * ((2 * i < N) ? i : i - N);
            acc.q0 += w;
            acc.q1 += k * w;
            acc.q2 += k * k * w;
        });
    }
    // Shear: the transformed line (pattern RL, bin i) times exp(-i k_i d) / N
    // (core/src/rotation.ixx:25-66), in FP64, k_i = dk i below N / 2 and
    // dk (i - N) above.  Along a butterfly's bins i = j + r q (q = N / RL) the
    // factor is exp(-i dk d j) exp(-i dk d q)^r, times exp(i dk d N) in the
    // upper half: three sin/cos per butterfly and a recurrence (relative
    // error below 1e-14) instead of a sin/cos per bin.
    __device__ __forceinline__ void shear_phase(double d) {
        constexpr int TT = N / E, Q = N / RL;
        const double base = -a.dk * d;
        double ss, cs, su, cu;
        sincos_small(base * Q, &ss, &cs);
        sincos_small(-base * N, &su, &cu);
#pragma unroll
        for (int b = 0; b < E / RL; ++b) {
            const int j = t + b * TT;
            double s0, c0;
            sincos_small(base * j, &s0, &c0);
            cx<double> w = {c0, s0};
#pragma unroll
            for (int r = 0; r < RL; ++r) {
                const int i = j + r * Q;
                const cx<double> f = 2 * i < N ? w : cmul(w, cx<double>{cu, su});
                v[b * RL + r] = mul_d(v[b * RL + r], cx<double>{f.x * a.scale, f.y * a.scale});
                w = cmul(w, cx<double>{cs, ss});
            }
        }
    }
    // v *= m (absorbing mask only), pattern R1
    __device__ __forceinline__ void mask_only() {
        if (!valid || !a.mask) return;
        for_each_point<N, E, R1>(t, [&](int s, int, int j, int ro) {
            const uint32_t vo = (base + j * es) * RS, so = ro * es * RS;
            const T c = M.ld(vo, so);
            if constexpr (sizeof(T) == 4)
                v[s] = {__builtin_fmaf(-v[s].x, c, v[s].x), __builtin_fmaf(-v[s].y, c, v[s].y)};
            else
                v[s] = cscale(v[s], c);
        });
    }
    __device__ __forceinline__ void fft() { Transform<T, N, E, -1, 1, Fwd>::run(v, t, tw, x); }
    __device__ __forceinline__ void ifft() { Transform<T, N, E, +1, 1, Inv>::run(v, t, tw, x); }
    // Four-step twiddles exp(DIR 2 pi i m k / n_total) of the points in
    // pattern R, k their position along the line, m the line's index.  Along
    // a butterfly's points k = j + r N / R, so each is the butterfly's first
    // times exp(DIR 2 pi i m (N / R) / n_total)^r: two exactly reduced
    // sine/cosine pairs per butterfly and a recurrence (at most 31 FP64
    // products, relative error below 1e-14) instead of one pair per point.
    template <int R, int DIR>
    __device__ __forceinline__ void step_twiddles(int m) {
        constexpr int TT = N / E;
        const cx<double> wq = step_twiddle<DIR>(m, N / R, a.n_total);
#pragma unroll
        for (int b = 0; b < E / R; ++b) {
            cx<double> w = step_twiddle<DIR>(m, t + b * TT, a.n_total);
#pragma unroll
            for (int r = 0; r < R; ++r) {
                v[b * R + r] = mul_d(v[b * R + r], w);
                if (r + 1 < R) w = cmul(w, wq);
            }
        }
    }
    // four-step twiddle after the column FFT: exp(-2 pi i n1 k2 / N), k2 = position
    __device__ __forceinline__ void twiddle_out(int n1) { step_twiddles<RL, -1>(n1); }
};

// ---- imaginary-time norm ----------------------------------------------------------

// Sum and max over the block, every thread receiving the same values: a
// fixed xor-shuffle tree per wave64 (lane 0's result), then the waves in order.
__device__ __forceinline__ void block_sum_max(double& s, double& m) {
    __shared__ double red[2 * 16];
    for (int off = 32; off > 0; off >>= 1) {
        s += __shfl_xor(s, off, 64);
        const double o = __shfl_xor(m, off, 64);
        m = o > m ? o : m;
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int nw = (blockDim.x + 63) >> 6;
    __syncthreads();  // red may still be read by a previous use
    if (lane == 0) {
        red[2 * wave] = s;
        red[2 * wave + 1] = m;
    }
    __syncthreads();
    s = 0.0;
    m = 0.0;
    for (int w = 0; w < nw; ++w) {
        s += red[2 * w];
        m = red[2 * w + 1] > m ? red[2 * w + 1] : m;
    }
}

// A measuring sweep's per-block (sum |psi|^2, max |psi|^2) partials -> the
// step's norm, peak and the next measuring sweep's scale.  Wave 0 of the
// block does it with a fixed partition (lane l adds partials l, l + 64, ...
// in order) and a fixed xor-shuffle tree, so the result is the same bits
// whatever kernel runs it and whatever its block size (every block holds at
// least one full wave); the other waves skip it.
__device__ __forceinline__ void relax_finish_block(const double* partials, RelaxSlots* slots,
                                                   int parity, double cell) {
    if (threadIdx.x >= 64) return;
    const unsigned nb = slots->nblocks;
    double ts = 0.0, tm = 0.0;
#pragma unroll 4
    for (unsigned b = threadIdx.x; b < nb; b += 64) {
        ts += partials[2 * b];
        const double q = partials[2 * b + 1];
        tm = q > tm ? q : tm;
    }
    for (int off = 32; off > 0; off >>= 1) {
        ts += __shfl_xor(ts, off, 64);
        const double o = __shfl_xor(tm, off, 64);
        tm = o > tm ? o : tm;
    }
    if (threadIdx.x == 0) {
        const double norm = ts * cell;
        slots->scale[parity ^ 1] = norm > 0.0 ? 1.0 / sqrt(norm) : 1.0;
        slots->norm_sq = norm;
        slots->peak = tm;
    }
}

// Sweeps that can follow a measuring sweep and finish its norm.
constexpr bool mode_finishes(int mode) {
    return mode == S_FWD || mode == S_KIN || mode == C_KIN_TW;
}

// End of a measuring sweep, per block: this block's (sum, max) of |psi|^2
// (fixed-order block reduction) into partials[2 b], partials[2 b + 1]; block 0
// records the grid size.  launch_relax_finish, queued right after the sweep,
// adds the partials in block order (bitwise repeatable) and leaves scale,
// norm and peak for the next sweep.  (A last-block-done ticket here needed an
// agent-scope release/acquire per block, i.e. an L2 writeback and invalidate
// per block on this multi-L2 part: it tripled the sweep's time.)
__device__ __forceinline__ void relax_block_done(const SweepArgs& a, SweepAcc ra) {
    double s = ra.q0, m = ra.q1;
    block_sum_max(s, m);
    if (threadIdx.x == 0) {
        a.partials[2 * blockIdx.x] = s;
        a.partials[2 * blockIdx.x + 1] = m;
        if (blockIdx.x == 0) a.relax->nblocks = gridDim.x;
    }
}

// End of an observing sweep, per block: the three sums (fixed-order block
// reduction) into partials[3 * block + q]; a single-block pass adds them.
__device__ __forceinline__ void obs_block_done(const SweepArgs& a, SweepAcc acc) {
    __shared__ double red[3 * 16];
    double q[3] = {acc.q0, acc.q1, acc.q2};
    for (int c = 0; c < 3; ++c)
        for (int off = 32; off > 0; off >>= 1) q[c] += __shfl_xor(q[c], off, 64);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int nw = (blockDim.x + 63) >> 6;
    if (lane == 0)
        for (int c = 0; c < 3; ++c) red[3 * wave + c] = q[c];
    __syncthreads();
    if (threadIdx.x < 3) {
        double t = 0.0;
        for (int w = 0; w < nw; ++w) t += red[3 * w + threadIdx.x];
        a.partials[3 * blockIdx.x + threadIdx.x] = t;
    }
}

// Contiguous sweeps that read every twiddle from the full staged table: the
// FP32 two-transform kick sweep between steps (with the wave-local exchange:
// x sweep at 256^3 FP32 61.2 -> 57.1 us, same box).  Elsewhere the recurrence
// stays: in other sweeps the table's reads raised the register count past an
// occupancy step, and the FP64 x sweep measured slower with it (115.7 ->
// 120.2 us).
template <typename T, int N, int MODE, int V>
constexpr bool contig_tw_table() {
    return sizeof(T) == 4 && Shape<T, N, V>::TW_TABLE_C && base_mode(MODE) == C_KICK_MID;
}
template <typename T, int N, int MODE, int V>
constexpr int contig_twn() {
    return contig_tw_table<T, N, MODE, V>() ? N : Shape<T, N, V>::TWN;
}

// ---- per-tile bodies ---------------------------------------------------------------

// One tile of a contiguous sweep: LPB whole lines.  `lds` holds LPB padded
// lines, `tw` the staged FP64 twiddles of this line length.
template <typename T, int N, int MODE, int V = 0>
__device__ __forceinline__ void contig_tile(const SweepArgs& a, long long tile, cx<T>* lds,
                                            const cx<double>* tw, int tid, SweepAcc& ra) {
    using S = Shape<T, N, V>;
    constexpr int BM = base_mode(MODE);
    constexpr bool DRV = mode_driven(MODE);
    using XC = ContigX<T, contig_tw_table<T, N, MODE, V>()>;
    using LT = Line<T, N, XC, V>;
    constexpr int TT = S::TT, LPB = S::LPB, E = S::E;
    static_assert(64 % TT == 0, "a contiguous line must lie inside one wave");
    const int l = tid / TT, t = tid % TT;
    XC x{lds + l * S::LSTR};
    const Buf<cx<T>> P(a.psi, a.psi_bytes);
    const Buf<T> F(a.phi, a.real_bytes);
    const Buf<T> M(a.mask, a.mask ? a.mask_bytes : 0u);
    const long long line = tile * LPB + l;
    LT L{a, P, F, M, x, tw, t, static_cast<uint32_t>(line * N), 1u, line < a.nlines};
    if (L.valid && a.mask && a.mask_sep) {
        L.my = a.mask_yz[line % a.dn1];
        L.mz = a.mask_yz[a.dn1 + line / a.dn1];
    }
    if constexpr (DRV) {
        // x-line `line` of the grid: y index line % ny, z index line / ny
        const int yj = a.dn1 > 0 ? static_cast<int>(line % a.dn1) : 0;
        const int zk = a.dn1 > 0 ? static_cast<int>(line / a.dn1) : 0;
        L.du0 = a.du_c + a.du_y * yj + a.du_z * zk;
        L.du1 = a.du_x;
    }
    if constexpr (BM == C_KICK_FIRST) {
        L.template load<LT::R1>();
        L.template kick<DRV>(1.0, false, a.dg);
        L.fft();
        L.template store<LT::RL>();
    } else if constexpr (BM == C_KICK_MID) {
        L.template load<LT::RL>();
        L.ifft();
        L.load_kick(true);
        L.template apply_kick<DRV>(2.0, true, a.dg);
        L.fft();
        L.template store<LT::RL>();
    } else if constexpr (BM == C_KICK_LAST) {
        L.template load<LT::RL>();
        L.ifft();
        L.load_kick(true);
        L.template apply_kick<DRV>(1.0, true, a.dg);
        L.template store<LT::R1>();
    } else if constexpr (BM == C_FWD) {
        L.template load<LT::R1>();
        L.fft();
        L.template store<LT::RL>();
    } else if constexpr (BM == C_KIN_TW) {
        // row k2 = line; after the row FFT position k1 holds bin k2 + N2*k1
        const cx<double>* kt = static_cast<const cx<double>*>(a.kt[2]);
        L.template load<LT::R1>();
        L.fft();
        for_each_point<N, E, LT::RL>(t, [&](int s, int i, int, int) {
            const int j = static_cast<int>(line) + a.n_rows * i;
            L.v[s] = mul_d(L.v[s], cscale(kt[j], a.scale));
        });
        L.ifft();
        L.template step_twiddles<LT::R1, +1>(static_cast<int>(line));
        L.template store<LT::R1>();
    } else if constexpr (BM == C_SINGLE) {
        const cx<double>* kt = static_cast<const cx<double>*>(a.kt[2]);
        // the drive's merged angle between steps s and s + 1 is twice theta
        // at their common time (core/src/drive.ixx:52-66)
        constexpr bool drive = DRV;
        L.template load<LT::R1>();
        L.load_kick(true);  // phases and mask stay in registers for every step
        L.template apply_kick<DRV>(1.0, false, drive ? -drive_theta(a, a.dt0) : 0.0);
        for (int step = 0; step < a.nsteps; ++step) {
            L.fft();
            for_each_point<N, E, LT::RL>(t, [&](int s, int i, int, int) {
                L.v[s] = mul_d(L.v[s], cscale(kt[i], a.scale));
            });
            L.ifft();
            const bool last = step + 1 == a.nsteps;
            const double th = drive ? drive_theta(a, a.dt0 + (step + 1) * a.ddt) : 0.0;
            L.template apply_kick<DRV>(last ? 1.0 : 2.0, true, last ? -th : -2.0 * th);
        }
        L.template store<LT::R1>();
    } else if constexpr (BM == C_OBS) {
        L.template load<LT::R1>();
        L.fft();
        L.observe_spectrum(ra);
    } else if constexpr (BM == C_SHEAR) {
        const long long idx = a.sh_sel ? line / a.dn1 : line % a.dn1;
        L.template load<LT::R1>();
        L.fft();
        L.shear_phase(a.sh_d0 + a.sh_d1 * static_cast<double>(idx));
        L.ifft();
        L.mask_only();
        L.template store<LT::R1>();
    } else if constexpr (BM == C_SHEAR_KICK_FIRST) {
        const long long idx = a.sh_sel ? line / a.dn1 : line % a.dn1;
        L.template load<LT::R1>();
        L.fft();
        L.shear_phase(a.sh_d0 + a.sh_d1 * static_cast<double>(idx));
        L.ifft();
        L.template kick<DRV>(1.0, false, 0.0);
        L.fft();
        L.template store<LT::RL>();
    } else if constexpr (BM == C_KICK_LAST_SHEAR) {
        const long long idx = a.sh_sel ? line / a.dn1 : line % a.dn1;
        L.template load<LT::RL>();
        L.ifft();
        L.load_kick(false);
        L.template apply_kick<DRV>(1.0, false, 0.0);
        L.fft();
        L.shear_phase(a.sh_d0 + a.sh_d1 * static_cast<double>(idx));
        L.ifft();
        L.mask_only();
        L.template store<LT::R1>();
    } else if constexpr (BM == C_RELAX_FIRST) {
        L.template load<LT::R1>();
        L.relax_first();
        L.fft();
        L.template store<LT::RL>();
    } else if constexpr (BM == C_RELAX_MID || BM == C_RELAX_LAST) {
        const double c = a.relax->scale[a.parity];
        L.template load<LT::RL>();
        L.ifft();
        L.relax_measure(c, BM == C_RELAX_MID, ra);
        if constexpr (BM == C_RELAX_MID) {
            L.fft();
            L.template store<LT::RL>();
        } else {
            L.template store<LT::R1>();
        }
    } else if constexpr (BM == C_RELAX_SINGLE) {
        // the whole line is in this block: each step is renormalised exactly
        // (core/src/imaginary_time.ixx:43-53)
        const cx<double>* kt = static_cast<const cx<double>*>(a.kt[2]);
        L.template load<LT::R1>();
        L.load_kick(false);
        double e[E];
        for (int s = 0; s < E; ++s) e[s] = L.valid ? exp(static_cast<double>(L.kp[s])) : 0.0;
        for (int s = 0; s < E; ++s)
            L.v[s] = {static_cast<T>(e[s] * L.v[s].x), static_cast<T>(e[s] * L.v[s].y)};
        double norm = 0.0, peak = 0.0;
        for (int step = 0; step < a.nsteps; ++step) {
            L.fft();
            for_each_point<N, E, LT::RL>(t, [&](int s, int i, int, int) {
                L.v[s] = mul_d(L.v[s], cscale(kt[i], a.scale));
            });
            L.ifft();
            double sum = 0.0, pk = 0.0;
            for (int s = 0; s < E; ++s) {
                const double xr = e[s] * L.v[s].x, xi = e[s] * L.v[s].y;
                const double a2 = xr * xr + xi * xi;
                sum += a2;
                pk = a2 > pk ? a2 : pk;
                L.v[s] = {static_cast<T>(xr), static_cast<T>(xi)};
            }
            block_sum_max(sum, pk);
            norm = sum * a.cell;
            peak = pk;
            const double c = norm > 0.0 ? 1.0 / sqrt(norm) : 1.0;
            const bool last = step + 1 == a.nsteps;
            for (int s = 0; s < E; ++s) {
                const double f = last ? c : c * e[s];
                L.v[s] = {static_cast<T>(f * L.v[s].x), static_cast<T>(f * L.v[s].y)};
            }
        }
        if (a.relax && tid == 0 && a.nsteps > 0) {
            a.relax->norm_sq = norm;
            a.relax->peak = peak;
        }
        L.template store<LT::R1>();
    }
}

// threadIdx.x as a value the compiler cannot see through (the asm emits no
// instruction): read once per tile (or per phase of a multi-phase tile), it
// keeps thread-indexed LDS and global offsets where they are used instead of
// hoisted out of the enclosing loop and held in VGPRs across it.
__device__ __forceinline__ int opaque_tid() {
    int t = static_cast<int>(threadIdx.x);
    asm volatile("" : "+v"(t));
    return t;
}

// One tile of a strided sweep: TX consecutive lines.  `kp` holds the staged
// line-axis kinetic table (times 1/N) for S_KIN.
template <typename T, int N, int MODE, int V = 0>
__device__ __forceinline__ void strided_tile(const SweepArgs& a, long long tile, cx<T>* lds,
                                             const cx<double>* tw, const cx<double>* kp, int tid,
                                             SweepAcc& ra) {
    using S = Shape<T, N, V>;
    constexpr int BM = base_mode(MODE);
    constexpr bool DRV = mode_driven(MODE);
    constexpr int TX = S::TX, E = S::E;
    using LT = Line<T, N, StridedX<T, TX>, V>;
    const int xi = tid % TX, t = tid / TX;
    StridedX<T, TX> x{lds + xi};
    const Buf<cx<T>> P(a.psi, a.psi_bytes);
    const Buf<T> F(a.phi, a.real_bytes);
    const Buf<T> M(a.mask, a.mask ? a.mask_bytes : 0u);
    const int xtiles = (a.nx + TX - 1) / TX;
    const int w = static_cast<int>(tile / xtiles);
    const int xx = static_cast<int>(tile % xtiles) * TX + xi;
    LT L{a, P, F, M, x, tw, t, static_cast<uint32_t>(xx) + a.wstride * w, a.estride,
         xx < a.nx};
    if constexpr (DRV) {
        // four-step column xx: point p of the line is x index xx + n1 p
        L.du0 = a.du_c + a.du_x * xx;
        L.du1 = a.du_x * a.dn1;
    }
    if constexpr (BM == S_FWD) {
        L.template load<LT::R1>();
        L.fft();
        L.template store<LT::RL>();
    } else if constexpr (BM == S_INV) {
        L.template load<LT::RL>();
        L.ifft();
        L.template store<LT::R1>();
    } else if constexpr (BM == S_KIN) {
        cx<double> kxw = {1.0, 0.0};
        if (a.kt[0] && L.valid) kxw = static_cast<const cx<double>*>(a.kt[0])[xx];
        if (a.kt[1]) kxw = cmul(kxw, static_cast<const cx<double>*>(a.kt[1])[w]);
        L.template load<LT::R1>();
        L.fft();
        for_each_point<N, E, LT::RL>(t, [&](int s, int i, int, int) {
            L.v[s] = mul_d(L.v[s], cmul(kxw, kp[i]));
            __builtin_amdgcn_sched_barrier(0);
        });
        L.ifft();
        L.template store<LT::R1>();
    } else if constexpr (BM == S_KICK_FIRST_TW) {
        L.template load<LT::R1>();
        L.template kick<DRV>(1.0, false, a.dg);
        L.fft();
        L.twiddle_out(xx);
        L.template store<LT::RL>();
    } else if constexpr (BM == S_KICK_MID_TW) {
        L.template load<LT::RL>();
        L.ifft();
        L.load_kick(true);
        L.template apply_kick<DRV>(2.0, true, a.dg);
        L.fft();
        L.twiddle_out(xx);
        L.template store<LT::RL>();
    } else if constexpr (BM == S_KICK_LAST) {
        L.template load<LT::RL>();
        L.ifft();
        L.load_kick(true);
        L.template apply_kick<DRV>(1.0, true, a.dg);
        L.template store<LT::R1>();
    } else if constexpr (BM == S_FWD_TW) {
        L.template load<LT::R1>();
        L.fft();
        L.twiddle_out(xx);
        L.template store<LT::RL>();
    } else if constexpr (BM == S_OBS) {
        L.template load<LT::R1>();
        L.fft();
        L.observe_spectrum(ra);
    } else if constexpr (BM == S_SHEAR) {
        const int idx = a.sh_sel ? w : xx;
        L.template load<LT::R1>();
        L.fft();
        L.shear_phase(a.sh_d0 + a.sh_d1 * static_cast<double>(idx));
        L.ifft();
        L.mask_only();
        L.template store<LT::R1>();
    } else if constexpr (BM == S_RELAX_FIRST_TW) {
        L.template load<LT::R1>();
        L.relax_first();
        L.fft();
        L.twiddle_out(xx);
        L.template store<LT::RL>();
    } else if constexpr (BM == S_RELAX_MID_TW || BM == S_RELAX_LAST) {
        const double c = a.relax->scale[a.parity];
        L.template load<LT::RL>();
        L.ifft();
        L.relax_measure(c, BM == S_RELAX_MID_TW, ra);
        if constexpr (BM == S_RELAX_MID_TW) {
            L.fft();
            L.twiddle_out(xx);
            L.template store<LT::RL>();
        } else {
            L.template store<LT::R1>();
        }
    }
}

// XCD-aware tile order of strided sweeps.  Workgroups are dispatched
// round-robin over the 8 XCDs (workgroup i on XCD i mod 8), each with its own
// L2.  Consecutive x-tiles read neighbouring pieces of the same rows; the
// G = ROW / (TX * sizeof(cx<T>)) x-tiles that together cover ROW bytes of
// every row go to consecutive workgroups of one XCD: launch index
// L = 8 G q + 8 j + x (x the XCD, j < G) sweeps tile 8 G q + G x + j.  A
// bijection on every whole chunk of 8 G; the last partial chunk keeps the
// identity order.
//
// ROW is 128 bytes (one cache line: tiles narrower than a line, 1024-point
// lines, share it in one L2; against the identity order 1024^2 FP32 34.1 ->
// 28.2 us per step, 1024 x 1024 x 16 FP32 381 -> 232 us, FP64 745 -> 505 us)
// except where a wider run measured faster on one
// box (whole steps, two alternated runs): 512-point lines 1024 bytes (512^3
// FP32 2095 -> 2053 us, FP64 4009 -> 3813 us), 128-point FP32 lines and
// 256-point FP64 lines 512 bytes (128^3 FP32 47.4 -> 44.4 us, 256^3 FP64
// 369 -> 364 us).  Wider runs for other lengths were neutral or slower (128^3
// FP64 +10% at 512 bytes, 192^3 +1-3%).
constexpr int kXcds = 8;
template <typename T, int N, int TX>
constexpr int line_group() {
    constexpr bool f64 = sizeof(T) == 8;
    constexpr int row = N == 512                               ? 1024
                        : (N == 128 && !f64) || (N == 256 && f64) ? 512
                                                                 : 128;
    constexpr int w = TX * static_cast<int>(sizeof(cx<T>));
    return w >= row ? 1 : row / w;
}
template <int G>
__device__ __forceinline__ long long xcd_tile(long long l, long long ntiles) {
    if constexpr (G == 1) {
        return l;
    } else {
        constexpr long long C = static_cast<long long>(kXcds) * G;
        if (l >= (ntiles / C) * C) return l;
        const long long q = l / C, r = l % C;
        return q * C + (r % kXcds) * G + r / kXcds;
    }
}

template <typename T, int N>
__device__ __forceinline__ void stage_kinetic(cx<double>* kp, const SweepArgs& a, int tid,
                                              int nthreads) {
    const cx<double>* src = static_cast<const cx<double>*>(a.kt[2]);
    for (int i = tid; i < N; i += nthreads) kp[i] = cscale(src[i], a.scale);
}

// ---- single-plan kernels (grid-stride over tiles) -------------------------------------

template <typename T, int N, int MODE, int V = 0>
__global__ void __launch_bounds__((Shape<T, N, V>::BLOCK_C), TDSE_MIN_WAVES)
contig_sweep(SweepArgs a) {
    using S = Shape<T, N, V>;
    static_assert(!mode_finishes(MODE) || S::BLOCK_C >= 64, "a finishing sweep holds a full wave");
    __shared__ cx<T> lds[S::LDS_C];
    constexpr int TWN = contig_twn<T, N, MODE, V>();
    __shared__ cx<double> lds_tw[TWN];
    const int tid = threadIdx.x;
    if constexpr (mode_finishes(MODE)) {
        if (a.fin_slots && blockIdx.x == 0)
            relax_finish_block(a.fin_partials, a.fin_slots, a.fin_parity, a.fin_cell);
    }
    if constexpr (S::EXCHANGE) {
        stage_table<TWN>(lds_tw, a.tw, tid, S::BLOCK_C);
        __syncthreads();
    }
    if constexpr (MODE == C_RELAX_FIRST) {
        // a call's first measuring sweep reads scale[0]: start it at 1
        if (blockIdx.x == 0 && tid == 0) a.relax->scale[0] = 1.0;
    }
    SweepAcc ra;
    const long long ntiles = (a.nlines + S::LPB - 1) / S::LPB;
    for (long long tile = blockIdx.x; tile < ntiles; tile += gridDim.x)
        contig_tile<T, N, MODE, V>(a, tile, lds, lds_tw, opaque_tid(), ra);
    if constexpr (mode_measures(MODE)) relax_block_done(a, ra);
    if constexpr (mode_observes(MODE)) obs_block_done(a, ra);
}

template <typename T, int N, int MODE, int V = 0>
__global__ void __launch_bounds__((Shape<T, N, V>::BLOCK_S), TDSE_MIN_WAVES)
strided_sweep(SweepArgs a) {
    using S = Shape<T, N, V>;
    static_assert(!mode_finishes(MODE) || S::BLOCK_S >= 64, "a finishing sweep holds a full wave");
    constexpr bool KIN = MODE == S_KIN;
    __shared__ cx<T> lds[S::LDS_S];
    __shared__ cx<double> lds_tw[S::TWN];
    __shared__ cx<double> lds_kp[KIN ? N : 1];
    const int tid = threadIdx.x;
    if constexpr (mode_finishes(MODE)) {
        if (a.fin_slots && blockIdx.x == 0)
            relax_finish_block(a.fin_partials, a.fin_slots, a.fin_parity, a.fin_cell);
    }
    if constexpr (S::EXCHANGE) stage_table<S::TWN>(lds_tw, a.tw, tid, S::BLOCK_S);
    if constexpr (KIN) stage_kinetic<T, N>(lds_kp, a, tid, S::BLOCK_S);
    if constexpr (S::EXCHANGE || KIN) __syncthreads();
    if constexpr (MODE == S_RELAX_FIRST_TW) {
        if (blockIdx.x == 0 && tid == 0) a.relax->scale[0] = 1.0;
    }
    SweepAcc ra;
    const long long ntiles = static_cast<long long>((a.nx + S::TX - 1) / S::TX) * a.nw;
    constexpr int G = line_group<T, N, S::TX>();
    for (long long l = blockIdx.x; l < ntiles; l += gridDim.x)
        strided_tile<T, N, MODE, V>(a, xcd_tile<G>(l, ntiles), lds, lds_tw, lds_kp, opaque_tid(),
                                    ra);
    if constexpr (mode_measures(MODE)) relax_block_done(a, ra);
    if constexpr (mode_observes(MODE)) obs_block_done(a, ra);
}

// ---- batched kernels: one launch over the tiles of many plans ------------------------

// Plans that share a line length and mode are swept by one launch.  Item i
// owns global tiles [first_tile[i], first_tile[i+1]); the last entry of
// first_tile is the total.  All items' twiddle tables are equal (same N).
struct BatchTable {
    const SweepArgs* items;
    const long long* first_tile;  // nitems + 1 entries
    int nitems;
    int nsteps;  // steps of the one-block all-steps mode (C_SINGLE) in this call
};

__device__ __forceinline__ int find_item(const BatchTable& b, long long tile) {
    int lo = 0, hi = b.nitems - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (b.first_tile[mid] <= tile) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

template <typename T, int N, int MODE, int V = 0>
__global__ void __launch_bounds__((Shape<T, N, V>::BLOCK_C), TDSE_MIN_WAVES)
contig_sweep_batch(BatchTable b) {
    using S = Shape<T, N, V>;
    __shared__ cx<T> lds[S::LDS_C];
    constexpr int TWN = contig_twn<T, N, MODE, V>();
    __shared__ cx<double> lds_tw[TWN];
    const int tid = threadIdx.x;
    if constexpr (S::EXCHANGE) {
        stage_table<TWN>(lds_tw, b.items[0].tw, tid, S::BLOCK_C);
        __syncthreads();
    }
    // contiguous chunk of tiles per block: consecutive tiles mostly share a plan
    const long long total = b.first_tile[b.nitems];
    const long long chunk = (total + gridDim.x - 1) / gridDim.x;
    const long long g0 = blockIdx.x * chunk;
    const long long g1 = g0 + chunk < total ? g0 + chunk : total;
    // one search per block; consecutive tiles belong to the same or a later item
    int it = g0 < g1 ? find_item(b, g0) : 0;
    for (long long g = g0; g < g1; ++g) {
        while (g >= b.first_tile[it + 1]) ++it;
        SweepAcc unused;
        if constexpr (MODE == C_SINGLE) {
            SweepArgs a = b.items[it];
            a.nsteps = b.nsteps;
            contig_tile<T, N, MODE, V>(a, g - b.first_tile[it], lds, lds_tw, opaque_tid(), unused);
        } else {
            contig_tile<T, N, MODE, V>(b.items[it], g - b.first_tile[it], lds, lds_tw,
                                       opaque_tid(), unused);
        }
    }
}

template <typename T, int N, int MODE, int V = 0>
__global__ void __launch_bounds__((Shape<T, N, V>::BLOCK_S), TDSE_MIN_WAVES)
strided_sweep_batch(BatchTable b) {
    using S = Shape<T, N, V>;
    constexpr bool KIN = MODE == S_KIN;
    __shared__ cx<T> lds[S::LDS_S];
    __shared__ cx<double> lds_tw[S::TWN];
    __shared__ cx<double> lds_kp[KIN ? N : 1];
    const int tid = threadIdx.x;
    if constexpr (S::EXCHANGE) {
        stage_table<S::TWN>(lds_tw, b.items[0].tw, tid, S::BLOCK_S);
        __syncthreads();
    }
    const long long total = b.first_tile[b.nitems];
    const long long chunk = (total + gridDim.x - 1) / gridDim.x;
    // one tile per workgroup: the XCD-aware order (a block of several
    // consecutive tiles already keeps a line group in one L2)
    const long long g0 = chunk == 1 ? xcd_tile<line_group<T, N, S::TX>()>(blockIdx.x, total)
                                    : blockIdx.x * chunk;
    const long long g1 = g0 + chunk < total ? g0 + chunk : total;
    int staged = -1;
    int it = g0 < g1 ? find_item(b, g0) : 0;
    for (long long g = g0; g < g1; ++g) {
        while (g >= b.first_tile[it + 1]) ++it;
        if constexpr (KIN) {
            if (it != staged) {  // block-uniform: every thread sees the same tile
                __syncthreads();
                stage_kinetic<T, N>(lds_kp, b.items[it], tid, S::BLOCK_S);
                __syncthreads();
                staged = it;
            }
        }
        SweepAcc unused;
        strided_tile<T, N, MODE, V>(b.items[it], g - b.first_tile[it], lds, lds_tw, lds_kp,
                                    opaque_tid(), unused);
    }
}

}  // namespace tdse
