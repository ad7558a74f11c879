// line_plans.hpp -- compile-time line FFT plans and launch shapes per length.
//
// For each supported line length N: E points per thread (T = N / E threads
// per line) and the forward radix list (product N, every radix divides E).
// The inverse transform runs the same radices in reverse order.
//
// Launch shapes:
//   contiguous sweeps (lines along the unit-stride axis): LPB lines per block;
//   strided sweeps (lines along a strided axis): TX consecutive unit-stride
//   positions per block, lane index x-fastest so every global access of a
//   wave covers TX consecutive elements.
#pragma once

#include "fft_engine.hpp"

namespace tdse {

template <int N>
struct LinePlan;

#define TDSE_LINE_PLAN(N_, E_, ...)        \
    template <>                            \
    struct LinePlan<N_> {                  \
        static constexpr int N = N_;       \
        static constexpr int E = E_;       \
        using Fwd = Radices<__VA_ARGS__>;  \
    };

TDSE_LINE_PLAN(2, 2, 2)
TDSE_LINE_PLAN(4, 4, 4)
TDSE_LINE_PLAN(8, 8, 8)
TDSE_LINE_PLAN(16, 16, 16)
TDSE_LINE_PLAN(32, 8, 8, 4)
TDSE_LINE_PLAN(64, 8, 8, 8)
TDSE_LINE_PLAN(128, 16, 16, 8)
TDSE_LINE_PLAN(256, 16, 16, 16)
TDSE_LINE_PLAN(512, 16, 8, 8, 8)
TDSE_LINE_PLAN(1024, 16, 16, 16, 4)
TDSE_LINE_PLAN(96, 12, 6, 4, 4)
TDSE_LINE_PLAN(160, 20, 10, 4, 4)
TDSE_LINE_PLAN(192, 24, 8, 6, 4)
#undef TDSE_LINE_PLAN

// X-macro over every compiled line length.
#define TDSE_FOR_EACH_N(X) \
    X(2) X(4) X(8) X(16) X(32) X(64) X(128) X(256) X(512) X(1024) X(96) X(160) X(192)

constexpr bool line_supported(int n) {
    switch (n) {
#define TDSE_CASE(N_) case N_:
        TDSE_FOR_EACH_N(TDSE_CASE)
#undef TDSE_CASE
        return true;
        default:
            return false;
    }
}

// LDS index padding: one spare slot every 16 points breaks the power-of-two
// strides of the Stockham scatter across LDS banks.
__device__ __forceinline__ constexpr int lds_pad(int i) { return i + (i >> 4); }
constexpr int lds_len(int n) { return n + ((n - 1) >> 4) + 1; }

constexpr int cmin(int a, int b) { return a < b ? a : b; }
constexpr int cmax(int a, int b) { return a > b ? a : b; }
constexpr int pow2_floor(int a) {
    int p = 1;
    while (2 * p <= a) p *= 2;
    return p;
}

constexpr int kLdsCap = 80 * 1024;

template <typename T, int N>
struct Shape {
    using P = LinePlan<N>;
    static constexpr int E = P::E;
    static constexpr int TT = N / E;  // threads per line
    static constexpr int PASSES = P::Fwd::count;
    static constexpr bool EXCHANGE = PASSES > 1;
    static constexpr int CSZ = static_cast<int>(sizeof(cx<T>));
    static constexpr int LSTR = lds_len(N);
    // FP64 tables staged in LDS: twiddles and (strided kinetic sweep) K_p
    static constexpr int DATA_CAP = (kLdsCap - 2 * N * 16) / CSZ;

    // contiguous sweeps
    static constexpr int LPB_WANT = cmax(1, 256 / TT);
    static constexpr int LPB_LDS = EXCHANGE ? cmax(1, DATA_CAP / LSTR) : LPB_WANT;
    static constexpr int LPB = pow2_floor(cmin(LPB_WANT, LPB_LDS));
    static constexpr int BLOCK_C = LPB * TT;
    static constexpr int LDS_C = EXCHANGE ? LPB * LSTR : 1;

    // strided sweeps
    static constexpr int TX_WANT = cmax(1, cmin(64, 256 / TT));
    static constexpr int TX_LDS = EXCHANGE ? cmax(1, DATA_CAP / LSTR) : TX_WANT;
    static constexpr int TX = pow2_floor(cmin(TX_WANT, TX_LDS));
    static constexpr int BLOCK_S = TX * TT;
    static constexpr int LDS_S = EXCHANGE ? TX * LSTR : 1;
};

}  // namespace tdse
