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
// Phases:  the kick reads phi = -V dt / 2 per cell (FP32 path: reduced modulo
// 2 pi in FP64 on the host), the kinetic angle -0.5 |k|^2 / m dt is formed in
// FP64 from the wavenumber index (FP32 path: reduced modulo 2 pi in FP64 on
// the device before the single-precision sincos).
#pragma once

#include "line_plans.hpp"

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
    kNumModes = 13
};

struct SweepArgs {
    void* psi;            // cx<T>*: the state, x-fastest
    const void* phi;      // T*: half-kick angle -V dt/2 per cell
    const void* mask;     // T*: per-step absorbing mask, or null
    const void* tw;       // cx<T>*: W_N^m = exp(-2 pi i m / N), m < N, for this line length
    long long nlines;     // contiguous: number of lines
    long long wstride;    // strided: element stride of the outer index w
    long long estride;    // strided: element stride along the line
    int nx;               // strided: extent of the unit-stride index x
    int nw;               // strided: extent of the outer index w
    // kinetic angle: k_a = dk[a] * shifted(idx_a, kn[a]) for slots (x, w, p)
    double dk[3];
    int kn[3];
    double dt;
    double mass;
    double scale;         // 1 / total points (inverse normalisation)
    int n_total;          // four-step twiddle modulus (1D), else 0
    int n_rows;           // four-step: number of rows N2 (contiguous K mapping)
    int nsteps;           // C_SINGLE: steps to run
};

// ---- phase helpers ------------------------------------------------------------

__device__ __forceinline__ int shifted(int j, int n) { return (2 * j < n) ? j : j - n; }

__device__ __forceinline__ double kin_angle(const SweepArgs& a, int ix, int iw, int ip) {
    const double kx = a.dk[0] * static_cast<double>(shifted(ix, a.kn[0]));
    const double kw = a.dk[1] * static_cast<double>(shifted(iw, a.kn[1]));
    const double kp = a.dk[2] * static_cast<double>(shifted(ip, a.kn[2]));
    return -0.5 * (kx * kx + kw * kw + kp * kp) / a.mass * a.dt;
}

// 2 pi split for FP64 argument reduction (hi has 33 significant bits, so
// n * hi is exact for |n| < 2^20).
constexpr double kTwoPiHi = 6.28318530717958623200e+00;
constexpr double kTwoPiLo = 2.44929359829470635445e-16;
constexpr double kInvTwoPi = 0.15915494309189533577;

__device__ __forceinline__ double reduce_2pi(double th) {
    const double n = rint(th * kInvTwoPi);
    return (th - n * kTwoPiHi) - n * kTwoPiLo;
}

template <typename T>
__device__ __forceinline__ cx<T> cis_scaled(double th, T scale);

template <>
__device__ __forceinline__ cx<double> cis_scaled<double>(double th, double scale) {
    double s, c;
    sincos(th, &s, &c);
    return {c * scale, s * scale};
}

template <>
__device__ __forceinline__ cx<float> cis_scaled<float>(double th, float scale) {
    const float r = static_cast<float>(reduce_2pi(th));
    float s, c;
    sincosf(r, &s, &c);
    return {c * scale, s * scale};
}

// kick factor m * exp(i * f * phi); f = 1 (half kick) or 2 (merged full kick)
template <typename T>
__device__ __forceinline__ cx<T> kick_factor(T phi, T f, T m) {
    T s, c;
    if constexpr (sizeof(T) == 8) {
        sincos(f * phi, &s, &c);
    } else {
        sincosf(f * phi, &s, &c);
    }
    return {c * m, s * m};
}

// four-step twiddle exp(DIR * 2 pi i (a*b mod n) / n), FP64 then rounded
template <typename T, int DIR>
__device__ __forceinline__ cx<T> step_twiddle(long long a, long long b, int n) {
    const long long m = (a * b) % n;
    double s, c;
    sincospi(2.0 * static_cast<double>(m) / static_cast<double>(n), &s, &c);
    return {static_cast<T>(c), static_cast<T>(DIR * s)};
}

// ---- LDS exchange buffers -------------------------------------------------------

template <typename T>
struct ContigX {
    cx<T>* base;  // this thread's line
    __device__ __forceinline__ void sync() const { __syncthreads(); }
    __device__ __forceinline__ void put(int i, cx<T> z) const { base[lds_pad(i)] = z; }
    __device__ __forceinline__ cx<T> get(int i) const { return base[lds_pad(i)]; }
};

template <typename T, int TX>
struct StridedX {
    cx<T>* base;  // lds + xi
    __device__ __forceinline__ void sync() const { __syncthreads(); }
    __device__ __forceinline__ void put(int i, cx<T> z) const { base[lds_pad(i) * TX] = z; }
    __device__ __forceinline__ cx<T> get(int i) const { return base[lds_pad(i) * TX]; }
};

template <typename T, int N, int SLOTS>
__device__ __forceinline__ const cx<T>* stage_twiddles(cx<T>* lds_tw, const void* tw_global,
                                                       int tid, int nthreads) {
    if constexpr (SLOTS > 0) {
        const cx<T>* g = static_cast<const cx<T>*>(tw_global);
        for (int i = tid; i < N; i += nthreads) lds_tw[i] = g[i];
        __syncthreads();
    }
    return lds_tw;
}

// ---- contiguous sweep -------------------------------------------------------------

template <typename T, int N, int MODE>
__global__ void __launch_bounds__((Shape<T, N>::BLOCK_C))
contig_sweep(SweepArgs a) {
    using S = Shape<T, N>;
    using Fwd = typename S::P::Fwd;
    using Inv = Reverse<Fwd>;
    constexpr int E = S::E, TT = S::TT, LPB = S::LPB;
    constexpr int R1 = First<Fwd>::value, RL = Last<Fwd>::value;

    __shared__ cx<T> lds[S::LDS_C];
    const int tid = threadIdx.x;
    const int l = tid / TT, t = tid % TT;
    const cx<T>* tw = stage_twiddles<T, N, S::TW_SLOTS>(lds + LPB * S::LSTR, a.tw, tid, S::BLOCK_C);
    ContigX<T> x{lds + l * S::LSTR};

    cx<T>* psi = static_cast<cx<T>*>(a.psi);
    const T* phi = static_cast<const T*>(a.phi);
    const T* mask = static_cast<const T*>(a.mask);
    const long long ntiles = (a.nlines + LPB - 1) / LPB;

    for (long long tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        const long long line = tile * LPB + l;
        const bool valid = line < a.nlines;
        const long long base = line * N;
        cx<T> v[E];
        auto load = [&](int i) { return valid ? psi[base + i] : cx<T>{0, 0}; };
        auto store = [&](int s, int i) {
            if (valid) psi[base + i] = v[s];
        };
        auto kick = [&](int s, int i, T f, bool with_mask) {
            if (!valid) return;
            const T m = (with_mask && mask) ? mask[base + i] : T(1);
            v[s] = cmul(v[s], kick_factor<T>(phi[base + i], f, m));
        };

        if constexpr (MODE == C_KICK_FIRST) {
            pass_gather<T, N, E, R1>(v, t, load);
            for_each_point<N, E, R1>(t, [&](int s, int i) { kick(s, i, T(1), false); });
            Transform<T, N, E, -1, 1, Fwd>::run(v, t, tw, x);
            for_each_point<N, E, RL>(t, store);
        } else if constexpr (MODE == C_KICK_MID) {
            pass_gather<T, N, E, RL>(v, t, load);
            Transform<T, N, E, +1, 1, Inv>::run(v, t, tw, x);
            for_each_point<N, E, R1>(t, [&](int s, int i) { kick(s, i, T(2), true); });
            Transform<T, N, E, -1, 1, Fwd>::run(v, t, tw, x);
            for_each_point<N, E, RL>(t, store);
        } else if constexpr (MODE == C_KICK_LAST) {
            pass_gather<T, N, E, RL>(v, t, load);
            Transform<T, N, E, +1, 1, Inv>::run(v, t, tw, x);
            for_each_point<N, E, R1>(t, [&](int s, int i) {
                kick(s, i, T(1), true);
                store(s, i);
            });
        } else if constexpr (MODE == C_KIN_TW) {
            // row k2 = line, position k1 holds X[k2 + N2*k1]
            pass_gather<T, N, E, R1>(v, t, load);
            Transform<T, N, E, -1, 1, Fwd>::run(v, t, tw, x);
            const T scale = static_cast<T>(a.scale);
            for_each_point<N, E, RL>(t, [&](int s, int i) {
                const int j = static_cast<int>(line) + a.n_rows * i;
                v[s] = cmul(v[s], cis_scaled<T>(kin_angle(a, 0, 0, j), scale));
            });
            Transform<T, N, E, +1, 1, Inv>::run(v, t, tw, x);
            for_each_point<N, E, R1>(t, [&](int s, int i) {
                v[s] = cmul(v[s], step_twiddle<T, +1>(i, line, a.n_total));
                store(s, i);
            });
        } else if constexpr (MODE == C_FWD) {
            pass_gather<T, N, E, R1>(v, t, load);
            Transform<T, N, E, -1, 1, Fwd>::run(v, t, tw, x);
            for_each_point<N, E, RL>(t, store);
        } else if constexpr (MODE == C_SINGLE) {
            pass_gather<T, N, E, R1>(v, t, load);
            // per-thread phases and mask stay in registers for the whole run
            cx<T> half[E], full[E];
            const T scale = static_cast<T>(a.scale);
            for_each_point<N, E, R1>(t, [&](int s, int i) {
                const T p = valid ? phi[base + i] : T(0);
                const T m = (valid && mask) ? mask[base + i] : T(1);
                half[s] = kick_factor<T>(p, T(1), T(1));
                full[s] = kick_factor<T>(p, T(2), m);
            });
            cx<T> kin[E];
            for_each_point<N, E, RL>(t, [&](int s, int i) {
                kin[s] = cis_scaled<T>(kin_angle(a, 0, 0, i), scale);
            });
#pragma unroll
            for (int s = 0; s < E; ++s) v[s] = cmul(v[s], half[s]);
            for (int step = 0; step < a.nsteps; ++step) {
                Transform<T, N, E, -1, 1, Fwd>::run(v, t, tw, x);
#pragma unroll
                for (int s = 0; s < E; ++s) v[s] = cmul(v[s], kin[s]);
                Transform<T, N, E, +1, 1, Inv>::run(v, t, tw, x);
                if (step + 1 < a.nsteps) {
#pragma unroll
                    for (int s = 0; s < E; ++s) v[s] = cmul(v[s], full[s]);
                } else {
                    for_each_point<N, E, R1>(t, [&](int s, int i) {
                        const T m = (valid && mask) ? mask[base + i] : T(1);
                        v[s] = cscale(cmul(v[s], half[s]), m);
                    });
                }
            }
            for_each_point<N, E, R1>(t, store);
        }
    }
}

// ---- strided sweep ------------------------------------------------------------------

template <typename T, int N, int MODE>
__global__ void __launch_bounds__((Shape<T, N>::BLOCK_S))
strided_sweep(SweepArgs a) {
    using S = Shape<T, N>;
    using Fwd = typename S::P::Fwd;
    using Inv = Reverse<Fwd>;
    constexpr int E = S::E, TX = S::TX;
    constexpr int R1 = First<Fwd>::value, RL = Last<Fwd>::value;

    __shared__ cx<T> lds[S::LDS_S];
    const int tid = threadIdx.x;
    const int xi = tid % TX, t = tid / TX;
    const cx<T>* tw = stage_twiddles<T, N, S::TW_SLOTS>(lds + TX * S::LSTR, a.tw, tid, S::BLOCK_S);
    StridedX<T, TX> x{lds + xi};

    cx<T>* psi = static_cast<cx<T>*>(a.psi);
    const T* phi = static_cast<const T*>(a.phi);
    const T* mask = static_cast<const T*>(a.mask);
    const int xtiles = (a.nx + TX - 1) / TX;
    const long long ntiles = static_cast<long long>(xtiles) * a.nw;

    for (long long tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        const int w = static_cast<int>(tile / xtiles);
        const int xx = static_cast<int>(tile % xtiles) * TX + xi;
        const bool valid = xx < a.nx;
        const long long base = xx + a.wstride * w;
        cx<T> v[E];
        auto addr = [&](int i) { return base + a.estride * i; };
        auto load = [&](int i) { return valid ? psi[addr(i)] : cx<T>{0, 0}; };
        auto store = [&](int s, int i) {
            if (valid) psi[addr(i)] = v[s];
        };
        auto kick = [&](int s, int i, T f, bool with_mask) {
            if (!valid) return;
            const long long g = addr(i);
            const T m = (with_mask && mask) ? mask[g] : T(1);
            v[s] = cmul(v[s], kick_factor<T>(phi[g], f, m));
        };
        auto twiddle_out = [&](int s, int i) {
            v[s] = cmul(v[s], step_twiddle<T, -1>(xx, i, a.n_total));
        };

        if constexpr (MODE == S_FWD) {
            pass_gather<T, N, E, R1>(v, t, load);
            Transform<T, N, E, -1, 1, Fwd>::run(v, t, tw, x);
            for_each_point<N, E, RL>(t, store);
        } else if constexpr (MODE == S_INV) {
            pass_gather<T, N, E, RL>(v, t, load);
            Transform<T, N, E, +1, 1, Inv>::run(v, t, tw, x);
            for_each_point<N, E, R1>(t, store);
        } else if constexpr (MODE == S_KIN) {
            pass_gather<T, N, E, R1>(v, t, load);
            Transform<T, N, E, -1, 1, Fwd>::run(v, t, tw, x);
            const T scale = static_cast<T>(a.scale);
            for_each_point<N, E, RL>(t, [&](int s, int i) {
                v[s] = cmul(v[s], cis_scaled<T>(kin_angle(a, xx, w, i), scale));
            });
            Transform<T, N, E, +1, 1, Inv>::run(v, t, tw, x);
            for_each_point<N, E, R1>(t, store);
        } else if constexpr (MODE == S_KICK_FIRST_TW) {
            pass_gather<T, N, E, R1>(v, t, load);
            for_each_point<N, E, R1>(t, [&](int s, int i) { kick(s, i, T(1), false); });
            Transform<T, N, E, -1, 1, Fwd>::run(v, t, tw, x);
            for_each_point<N, E, RL>(t, [&](int s, int i) {
                twiddle_out(s, i);
                store(s, i);
            });
        } else if constexpr (MODE == S_KICK_MID_TW) {
            pass_gather<T, N, E, RL>(v, t, load);
            Transform<T, N, E, +1, 1, Inv>::run(v, t, tw, x);
            for_each_point<N, E, R1>(t, [&](int s, int i) { kick(s, i, T(2), true); });
            Transform<T, N, E, -1, 1, Fwd>::run(v, t, tw, x);
            for_each_point<N, E, RL>(t, [&](int s, int i) {
                twiddle_out(s, i);
                store(s, i);
            });
        } else if constexpr (MODE == S_FWD_TW) {
            pass_gather<T, N, E, R1>(v, t, load);
            Transform<T, N, E, -1, 1, Fwd>::run(v, t, tw, x);
            for_each_point<N, E, RL>(t, [&](int s, int i) {
                twiddle_out(s, i);
                store(s, i);
            });
        } else if constexpr (MODE == S_KICK_LAST) {
            pass_gather<T, N, E, RL>(v, t, load);
            Transform<T, N, E, +1, 1, Inv>::run(v, t, tw, x);
            for_each_point<N, E, R1>(t, [&](int s, int i) {
                kick(s, i, T(1), true);
                store(s, i);
            });
        }
    }
}

}  // namespace tdse
