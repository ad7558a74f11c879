// line_plans.hpp -- compile-time line FFT plans and launch shapes per length.
//
// For each supported line length N: E points per thread (T = N / E threads
// per line) and the forward radix list (product N, every radix divides E).
// The inverse transform runs the same radices in reverse order.  ALT = 1
// selects an alternative radix list with a different E (fewer or more points
// per thread); lengths without one reuse the base plan.
//
// A kernel variant V (0..3) combines the radix list (V & 1) with a tile-width
// multiplier (V & 2: twice the lines per block).  The autotuner times the
// variants of each sweep and keeps the fastest.
//
// Launch shapes:
//   contiguous sweeps (lines along the unit-stride axis): LPB lines per block;
//   strided sweeps (lines along a strided axis): TX consecutive unit-stride
//   positions per block, lane index x-fastest so every global access of a
//   wave covers TX consecutive elements.
#pragma once

#include "fft_engine.hpp"

namespace tdse {

template <int N, int ALT = 0>
struct LinePlan;

#define TDSE_LINE_PLAN(N_, ALT_, E_, ...)  \
    template <>                            \
    struct LinePlan<N_, ALT_> {            \
        static constexpr int N = N_;       \
        static constexpr int E = E_;       \
        using Fwd = Radices<__VA_ARGS__>;  \
    };

TDSE_LINE_PLAN(2, 0, 2, 2)
TDSE_LINE_PLAN(4, 0, 4, 4)
TDSE_LINE_PLAN(8, 0, 8, 8)
TDSE_LINE_PLAN(16, 0, 16, 16)
TDSE_LINE_PLAN(32, 0, 8, 8, 4)
TDSE_LINE_PLAN(64, 0, 8, 8, 8)
TDSE_LINE_PLAN(128, 0, 16, 16, 8)
TDSE_LINE_PLAN(256, 0, 16, 16, 16)
TDSE_LINE_PLAN(512, 0, 16, 8, 8, 8)
TDSE_LINE_PLAN(1024, 0, 16, 16, 16, 4)
TDSE_LINE_PLAN(96, 0, 12, 6, 4, 4)
TDSE_LINE_PLAN(160, 0, 20, 10, 4, 4)
TDSE_LINE_PLAN(192, 0, 12, 12, 4, 4)
// alternatives
TDSE_LINE_PLAN(64, 1, 4, 4, 4, 4)
TDSE_LINE_PLAN(128, 1, 8, 8, 4, 4)
TDSE_LINE_PLAN(256, 1, 8, 8, 8, 4)
TDSE_LINE_PLAN(512, 1, 8, 8, 8, 8)
TDSE_LINE_PLAN(1024, 1, 32, 32, 32)
TDSE_LINE_PLAN(96, 1, 24, 12, 8)
TDSE_LINE_PLAN(192, 1, 24, 8, 6, 4)
#undef TDSE_LINE_PLAN

template <int N>
struct LinePlan<N, 1> : LinePlan<N, 0> {};

// X-macro over every compiled line length.
#define TDSE_FOR_EACH_N(X) \
    X(2) X(4) X(8) X(16) X(32) X(64) X(128) X(256) X(512) X(1024) X(96) X(160) X(192)
// Lengths whose variants 1..3 are compiled (the others run variant 0 only).
#define TDSE_FOR_EACH_TUNED_N(X) X(64) X(128) X(256) X(512) X(1024) X(96) X(192)
constexpr int kVariants = 4;

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
// Cap (bytes) on the line-exchange LDS of one block, so that at least four
// blocks share a CU's 160 KB: fewer lines (or x columns) per block where the
// lines are long.  Measured on one box (variant 0, HIP events): 160^3 FP32
// 86 -> 81 us per step, 192^3 FP32 122 -> 116, 192^3 FP64 206 -> 187, 256^3
// FP64 379 -> 370, 256^3 FP32 unchanged (its tiles already fit).  0 = no cap.
#ifndef TDSE_TILE_LDS
#define TDSE_TILE_LDS 40960
#endif
// Narrow strided tiles (rows under 128 bytes: lines of 512 and more) may take
// up to this much LDS, tables included, to reach 128-byte rows (two blocks
// per CU).  Measured (variant 0, one box): 512^3 FP32 2287 -> 2086 us per
// step (z sweep 745 -> 609 us), 512^3 FP64 4537 -> 4013; 256^3 unchanged.
// Rows of 256 bytes (160 KB of LDS, one block per CU) measured slower:
// 512^3 FP32 z sweep 610 -> 797 us, FP64 step 4009 -> 3979 us.  0 = off.
#ifndef TDSE_NARROW_LDS
#define TDSE_NARROW_LDS 81920
#endif

template <typename T, int N, int V = 0>
struct Shape {
    using P = LinePlan<N, V & 1>;
    static constexpr int WIDE = (V & 2) ? 2 : 1;  // tile-width multiplier
    static constexpr int E = P::E;
    static constexpr int TT = N / E;  // threads per line
    static constexpr int PASSES = P::Fwd::count;
    static constexpr bool EXCHANGE = PASSES > 1;
    // Twiddle entries the passes read: a pass of radix R after Ns points reads
    // W^{k N / (Ns R)}, k < Ns, all below N / R; the inverse runs the radix list
    // reversed, so every radix has a pass after the first: N / (least radix).
    static constexpr int TWN = EXCHANGE ? N / MinRadix<typename P::Fwd>::value : 1;
    // Contiguous sweeps of lines up to 256 points may stage the whole table
    // (4 KB at most: the tiles still fit four blocks per CU) and read every
    // twiddle from it, four FP64 operations fewer per twiddle than the
    // recurrence (which sweeps do: see contig_tw_table).
    static constexpr bool TW_TABLE_C = EXCHANGE && N <= 256;
    static constexpr int CSZ = static_cast<int>(sizeof(cx<T>));
    static constexpr int LSTR = lds_len(N);
    // FP64 tables staged in LDS: twiddles and (strided kinetic sweep) K_p
    static constexpr int DATA_CAP = (kLdsCap - 2 * N * 16) / CSZ;

    // contiguous sweeps
    static constexpr int LPB_WANT = cmax(1, WIDE * 256 / TT);
    static constexpr int LPB_LDS = EXCHANGE ? cmax(1, DATA_CAP / LSTR) : LPB_WANT;
    static constexpr int OCC_LINES =
        TDSE_TILE_LDS > 0 ? cmax(1, TDSE_TILE_LDS / (LSTR * CSZ)) : (1 << 30);
    static constexpr int LPB = pow2_floor(cmin(cmin(LPB_WANT, LPB_LDS), OCC_LINES));
    static constexpr int BLOCK_C = LPB * TT;
    static constexpr int LDS_C = EXCHANGE ? LPB * LSTR : 1;

    // strided sweeps
    static constexpr int TX_WANT = cmax(1, cmin(64, WIDE * 256 / TT));
    static constexpr int TX_LDS = EXCHANGE ? cmax(1, DATA_CAP / LSTR) : TX_WANT;
    static constexpr int TX_BASE = pow2_floor(cmin(cmin(TX_WANT, TX_LDS), OCC_LINES));
    // Narrow tiles (rows under 128 bytes: long lines) read rows of a few
    // dozen bytes far apart; they may take up to TDSE_NARROW_LDS of LDS
    // (tables included) to reach 128-byte rows.
    static constexpr int TX_ROW = 128 / CSZ;
    static constexpr int NARROW_TX =
        TDSE_NARROW_LDS > 0 && EXCHANGE
            ? cmax(1, (TDSE_NARROW_LDS - TWN * 16 - N * 16) / (LSTR * CSZ))
            : 0;
    static constexpr int TX = TX_BASE * CSZ >= 128 || NARROW_TX <= TX_BASE
                                  ? TX_BASE
                                  : pow2_floor(cmin(cmin(cmax(TX_WANT, TX_ROW), 1024 / TT), NARROW_TX));
    static constexpr int BLOCK_S = TX * TT;
    static constexpr int LDS_S = EXCHANGE ? TX * LSTR : 1;
};

}  // namespace tdse
