// kernels_impl.hpp -- explicit instantiation of every (line length, mode)
// sweep for one precision.  Included by kernels_f32.hip / kernels_f64.hip
// with TDSE_REAL and TDSE_SUFFIX defined.
#pragma once

#include "launch.hpp"

namespace tdse {
namespace {

template <typename T, int N, int MODE, int V = 0>
hipError_t launch_sweep(const SweepArgs& a, int max_blocks, hipStream_t stream) {
    using S = Shape<T, N, V>;
    long long tiles;
    int block;
    if constexpr (mode_contiguous(MODE)) {
        tiles = (a.nlines + S::LPB - 1) / S::LPB;
        block = S::BLOCK_C;
    } else {
        tiles = static_cast<long long>((a.nx + S::TX - 1) / S::TX) * a.nw;
        block = S::BLOCK_S;
    }
    if (tiles <= 0) return hipSuccess;
    // measuring sweeps keep one partial per block
    if (mode_measures(MODE) && max_blocks > relax_max_blocks<T>()) max_blocks = relax_max_blocks<T>();
    if (mode_observes(MODE) && max_blocks > kObsMaxBlocks) max_blocks = kObsMaxBlocks;
    const int grid = static_cast<int>(tiles < max_blocks ? tiles : max_blocks);
    const LaunchTiming tm = g_launch_timing;
    if (tm.start) {
        g_launch_timing = LaunchTiming{};
        if constexpr (mode_contiguous(MODE)) {
            hipExtLaunchKernelGGL((contig_sweep<T, N, MODE, V>), dim3(grid), dim3(block), 0, stream,
                                  tm.start, tm.stop, 0, a);
        } else {
            hipExtLaunchKernelGGL((strided_sweep<T, N, MODE, V>), dim3(grid), dim3(block), 0, stream,
                                  tm.start, tm.stop, 0, a);
        }
    } else if constexpr (mode_contiguous(MODE)) {
        hipLaunchKernelGGL((contig_sweep<T, N, MODE, V>), dim3(grid), dim3(block), 0, stream, a);
    } else {
        hipLaunchKernelGGL((strided_sweep<T, N, MODE, V>), dim3(grid), dim3(block), 0, stream, a);
    }
    return hipGetLastError();
}

template <typename T, int N, int MODE, int V = 0>
hipError_t launch_batch(const BatchTable& b, long long tiles, int max_blocks, hipStream_t stream) {
    using S = Shape<T, N, V>;
    if (tiles <= 0) return hipSuccess;
    const int grid = static_cast<int>(tiles < max_blocks ? tiles : max_blocks);
    if constexpr (mode_contiguous(MODE)) {
        hipLaunchKernelGGL((contig_sweep_batch<T, N, MODE, V>), dim3(grid), dim3(S::BLOCK_C), 0,
                           stream, b);
    } else {
        hipLaunchKernelGGL((strided_sweep_batch<T, N, MODE, V>), dim3(grid), dim3(S::BLOCK_S), 0,
                           stream, b);
    }
    return hipGetLastError();
}

// The batched modes; variants 1..3 for the modes that have them (C_SINGLE:
// variant 0 only).
template <typename T, int N, int V = 0>
BatchLaunchFn pick_batch_mode(int mode) {
    switch (mode) {
        case C_KICK_FIRST: return &launch_batch<T, N, C_KICK_FIRST, V>;
        case C_KICK_MID: return &launch_batch<T, N, C_KICK_MID, V>;
        case C_KICK_LAST: return &launch_batch<T, N, C_KICK_LAST, V>;
        case C_KIN_TW: return &launch_batch<T, N, C_KIN_TW, V>;
        case C_SINGLE:
            if constexpr (V == 0) return &launch_batch<T, N, C_SINGLE>;
            else return nullptr;
        case S_FWD: return &launch_batch<T, N, S_FWD, V>;
        case S_INV: return &launch_batch<T, N, S_INV, V>;
        case S_KIN: return &launch_batch<T, N, S_KIN, V>;
        case S_KICK_FIRST_TW: return &launch_batch<T, N, S_KICK_FIRST_TW, V>;
        case S_KICK_MID_TW: return &launch_batch<T, N, S_KICK_MID_TW, V>;
        case S_KICK_LAST: return &launch_batch<T, N, S_KICK_LAST, V>;
        default: return nullptr;
    }
}

template <typename T, int N>
LaunchFn pick_mode(int mode) {
    switch (mode) {
        case C_KICK_FIRST: return &launch_sweep<T, N, C_KICK_FIRST>;
        case C_KICK_MID: return &launch_sweep<T, N, C_KICK_MID>;
        case C_KICK_LAST: return &launch_sweep<T, N, C_KICK_LAST>;
        case C_KIN_TW: return &launch_sweep<T, N, C_KIN_TW>;
        case C_SINGLE: return &launch_sweep<T, N, C_SINGLE>;
        case C_FWD: return &launch_sweep<T, N, C_FWD>;
        case S_FWD: return &launch_sweep<T, N, S_FWD>;
        case S_INV: return &launch_sweep<T, N, S_INV>;
        case S_KIN: return &launch_sweep<T, N, S_KIN>;
        case S_KICK_FIRST_TW: return &launch_sweep<T, N, S_KICK_FIRST_TW>;
        case S_KICK_MID_TW: return &launch_sweep<T, N, S_KICK_MID_TW>;
        case S_KICK_LAST: return &launch_sweep<T, N, S_KICK_LAST>;
        case S_FWD_TW: return &launch_sweep<T, N, S_FWD_TW>;
        case C_RELAX_FIRST: return &launch_sweep<T, N, C_RELAX_FIRST>;
        case C_RELAX_MID: return &launch_sweep<T, N, C_RELAX_MID>;
        case C_RELAX_LAST: return &launch_sweep<T, N, C_RELAX_LAST>;
        case C_RELAX_SINGLE: return &launch_sweep<T, N, C_RELAX_SINGLE>;
        case S_RELAX_FIRST_TW: return &launch_sweep<T, N, S_RELAX_FIRST_TW>;
        case S_RELAX_MID_TW: return &launch_sweep<T, N, S_RELAX_MID_TW>;
        case S_RELAX_LAST: return &launch_sweep<T, N, S_RELAX_LAST>;
        case C_OBS: return &launch_sweep<T, N, C_OBS>;
        case S_OBS: return &launch_sweep<T, N, S_OBS>;
        case C_SHEAR: return &launch_sweep<T, N, C_SHEAR>;
        case S_SHEAR: return &launch_sweep<T, N, S_SHEAR>;
        case C_SHEAR_KICK_FIRST: return &launch_sweep<T, N, C_SHEAR_KICK_FIRST>;
        case C_KICK_LAST_SHEAR: return &launch_sweep<T, N, C_KICK_LAST_SHEAR>;
        case C_KICK_FIRST | kDrive: return &launch_sweep<T, N, C_KICK_FIRST | kDrive>;
        case C_KICK_MID | kDrive: return &launch_sweep<T, N, C_KICK_MID | kDrive>;
        case C_KICK_LAST | kDrive: return &launch_sweep<T, N, C_KICK_LAST | kDrive>;
        case C_SINGLE | kDrive: return &launch_sweep<T, N, C_SINGLE | kDrive>;
        case S_KICK_FIRST_TW | kDrive: return &launch_sweep<T, N, S_KICK_FIRST_TW | kDrive>;
        case S_KICK_MID_TW | kDrive: return &launch_sweep<T, N, S_KICK_MID_TW | kDrive>;
        case S_KICK_LAST | kDrive: return &launch_sweep<T, N, S_KICK_LAST | kDrive>;
        default: return nullptr;
    }
}

template <typename T, int N, int V = 0>
SweepShape shape_of(int mode) {
    using S = Shape<T, N, V>;
    if (mode_is_contiguous(mode)) return {S::BLOCK_C, S::LPB};
    return {S::BLOCK_S, S::TX};
}

// Variants 1..3 of the step-path modes (tuned lengths only).
template <typename T, int N, int V>
LaunchFn pick_variant_mode(int mode) {
    switch (mode) {
        case C_KICK_FIRST: return &launch_sweep<T, N, C_KICK_FIRST, V>;
        case C_KICK_MID: return &launch_sweep<T, N, C_KICK_MID, V>;
        case C_KICK_LAST: return &launch_sweep<T, N, C_KICK_LAST, V>;
        case C_KIN_TW: return &launch_sweep<T, N, C_KIN_TW, V>;
        case S_FWD: return &launch_sweep<T, N, S_FWD, V>;
        case S_INV: return &launch_sweep<T, N, S_INV, V>;
        case S_KIN: return &launch_sweep<T, N, S_KIN, V>;
        case S_KICK_FIRST_TW: return &launch_sweep<T, N, S_KICK_FIRST_TW, V>;
        case S_KICK_MID_TW: return &launch_sweep<T, N, S_KICK_MID_TW, V>;
        case S_KICK_LAST: return &launch_sweep<T, N, S_KICK_LAST, V>;
        default: return nullptr;
    }
}



}  // namespace

#define TDSE_CAT2(a, b) a##b
#define TDSE_CAT(a, b) TDSE_CAT2(a, b)

#if defined(TDSE_VARIANT_UNIT)
LaunchFn TDSE_CAT(TDSE_CAT(find_variant_launcher_, TDSE_SUFFIX), TDSE_CAT(_v, TDSE_VARIANT_UNIT))(
    int n, int mode) {
    switch (n) {
#define TDSE_CASE(N_) \
    case N_:          \
        return pick_variant_mode<TDSE_REAL, N_, TDSE_VARIANT_UNIT>(mode);
        TDSE_FOR_EACH_TUNED_N(TDSE_CASE)
#undef TDSE_CASE
        default:
            return nullptr;
    }
}
SweepShape TDSE_CAT(TDSE_CAT(sweep_shape_, TDSE_SUFFIX), TDSE_CAT(_v, TDSE_VARIANT_UNIT))(int n,
                                                                                   int mode) {
    switch (n) {
#define TDSE_CASE(N_) \
    case N_:          \
        return shape_of<TDSE_REAL, N_, TDSE_VARIANT_UNIT>(mode);
        TDSE_FOR_EACH_TUNED_N(TDSE_CASE)
#undef TDSE_CASE
        default:
            return {0, 0};
    }
}
BatchLaunchFn TDSE_CAT(TDSE_CAT(find_batch_variant_launcher_, TDSE_SUFFIX),
                       TDSE_CAT(_v, TDSE_VARIANT_UNIT))(int n, int mode) {
    switch (n) {
#define TDSE_CASE(N_) \
    case N_:          \
        return pick_batch_mode<TDSE_REAL, N_, TDSE_VARIANT_UNIT>(mode);
        TDSE_FOR_EACH_TUNED_N(TDSE_CASE)
#undef TDSE_CASE
        default:
            return nullptr;
    }
}
#elif !defined(TDSE_BATCH_UNIT)
LaunchFn TDSE_CAT(find_launcher_, TDSE_SUFFIX)(int n, int mode) {
    switch (n) {
#define TDSE_CASE(N_) \
    case N_:          \
        return pick_mode<TDSE_REAL, N_>(mode);
        TDSE_FOR_EACH_N(TDSE_CASE)
#undef TDSE_CASE
        default:
            return nullptr;
    }
}

#else
BatchLaunchFn TDSE_CAT(find_batch_launcher_, TDSE_SUFFIX)(int n, int mode) {
    switch (n) {
#define TDSE_CASE(N_) \
    case N_:          \
        return pick_batch_mode<TDSE_REAL, N_>(mode);
        TDSE_FOR_EACH_N(TDSE_CASE)
#undef TDSE_CASE
        default:
            return nullptr;
    }
}

#endif

#if !defined(TDSE_BATCH_UNIT) && !defined(TDSE_VARIANT_UNIT)
SweepShape TDSE_CAT(sweep_shape_, TDSE_SUFFIX)(int n, int mode) {
    switch (n) {
#define TDSE_CASE(N_) \
    case N_:          \
        return shape_of<TDSE_REAL, N_>(mode);
        TDSE_FOR_EACH_N(TDSE_CASE)
#undef TDSE_CASE
        default:
            return {0, 0};
    }
}

#endif

}  // namespace tdse
