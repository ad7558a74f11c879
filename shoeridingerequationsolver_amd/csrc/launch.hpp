// launch.hpp -- host-side view of the compiled sweep kernels.
#pragma once

#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include "sweeps.hpp"

namespace tdse {

// Launch one sweep over the whole array described by `a` on `stream`.
using LaunchFn = hipError_t (*)(const SweepArgs& a, int max_blocks, hipStream_t stream);

// Kernel-boundary timing (per-sweep profiling): when `start` is set, the next
// single-plan sweep launched from this thread records `start` and `stop` at
// its kernel's own begin and end (hipExtLaunchKernelGGL, no marker packets
// around it) and clears both.  A sweep with nothing to launch leaves them set.
struct LaunchTiming {
    hipEvent_t start = nullptr;
    hipEvent_t stop = nullptr;
};
extern thread_local LaunchTiming g_launch_timing;

struct SweepShape {
    int block;      // threads per block
    int per_block;  // lines (contiguous) or x positions (strided) per block
};

// Launch one batched sweep: the tiles of every plan in `b` (same line length
// and mode), `tiles` in total.
using BatchLaunchFn = hipError_t (*)(const BatchTable& b, long long tiles, int max_blocks,
                                     hipStream_t stream);

BatchLaunchFn find_batch_launcher_f32(int n, int mode);
BatchLaunchFn find_batch_launcher_f64(int n, int mode);

// Variants 1..3 of a (line length, mode) sweep, one translation unit each;
// null if not compiled.
LaunchFn find_variant_launcher_f32_v1(int n, int mode);
LaunchFn find_variant_launcher_f32_v2(int n, int mode);
LaunchFn find_variant_launcher_f32_v3(int n, int mode);
LaunchFn find_variant_launcher_f64_v1(int n, int mode);
LaunchFn find_variant_launcher_f64_v2(int n, int mode);
LaunchFn find_variant_launcher_f64_v3(int n, int mode);

BatchLaunchFn find_batch_variant_launcher_f32_v1(int n, int mode);
BatchLaunchFn find_batch_variant_launcher_f32_v2(int n, int mode);
BatchLaunchFn find_batch_variant_launcher_f32_v3(int n, int mode);
BatchLaunchFn find_batch_variant_launcher_f64_v1(int n, int mode);
BatchLaunchFn find_batch_variant_launcher_f64_v2(int n, int mode);
BatchLaunchFn find_batch_variant_launcher_f64_v3(int n, int mode);

SweepShape sweep_shape_f32_v1(int n, int mode);
SweepShape sweep_shape_f32_v2(int n, int mode);
SweepShape sweep_shape_f32_v3(int n, int mode);
SweepShape sweep_shape_f64_v1(int n, int mode);
SweepShape sweep_shape_f64_v2(int n, int mode);
SweepShape sweep_shape_f64_v3(int n, int mode);

// Batched launcher of (line length, mode) in kernel variant `variant`
// (0: the base kernels); null if not compiled.
inline BatchLaunchFn find_batch_launcher(bool fp64, int n, int mode, int variant) {
    switch (variant) {
        case 0: return fp64 ? find_batch_launcher_f64(n, mode) : find_batch_launcher_f32(n, mode);
        case 1: return fp64 ? find_batch_variant_launcher_f64_v1(n, mode)
                            : find_batch_variant_launcher_f32_v1(n, mode);
        case 2: return fp64 ? find_batch_variant_launcher_f64_v2(n, mode)
                            : find_batch_variant_launcher_f32_v2(n, mode);
        case 3: return fp64 ? find_batch_variant_launcher_f64_v3(n, mode)
                            : find_batch_variant_launcher_f32_v3(n, mode);
        default: return nullptr;
    }
}

SweepShape sweep_shape_f32(int n, int mode);
SweepShape sweep_shape_f64(int n, int mode);
// Block size and tile width of (line length, mode) in kernel variant `variant`.
inline SweepShape sweep_shape(bool fp64, int n, int mode, int variant) {
    switch (variant) {
        case 0: return fp64 ? sweep_shape_f64(n, mode) : sweep_shape_f32(n, mode);
        case 1: return fp64 ? sweep_shape_f64_v1(n, mode) : sweep_shape_f32_v1(n, mode);
        case 2: return fp64 ? sweep_shape_f64_v2(n, mode) : sweep_shape_f32_v2(n, mode);
        case 3: return fp64 ? sweep_shape_f64_v3(n, mode) : sweep_shape_f32_v3(n, mode);
        default: return {0, 0};
    }
}

inline LaunchFn find_variant_launcher_f32(int n, int mode, int variant) {
    switch (variant) {
        case 1: return find_variant_launcher_f32_v1(n, mode);
        case 2: return find_variant_launcher_f32_v2(n, mode);
        case 3: return find_variant_launcher_f32_v3(n, mode);
        default: return nullptr;
    }
}
inline LaunchFn find_variant_launcher_f64(int n, int mode, int variant) {
    switch (variant) {
        case 1: return find_variant_launcher_f64_v1(n, mode);
        case 2: return find_variant_launcher_f64_v2(n, mode);
        case 3: return find_variant_launcher_f64_v3(n, mode);
        default: return nullptr;
    }
}

// Lookup of the compiled kernel for (line length, mode); null if not compiled.
LaunchFn find_launcher_f32(int n, int mode);
LaunchFn find_launcher_f64(int n, int mode);
SweepShape sweep_shape_f32(int n, int mode);
SweepShape sweep_shape_f64(int n, int mode);

inline bool mode_is_contiguous(int mode) { return mode_contiguous(mode); }

// Set the calling thread's tdse_last_error text; returns status.
int record_error(int status, const char* what);

}  // namespace tdse
