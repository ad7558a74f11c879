// grid.hpp -- grid coordinates on the device.
#pragma once

#include <hip/hip_runtime.h>

namespace tdse {

// Grid coordinate lo + i h with the product and the sum each rounded (no
// fused multiply-add), as core/src/grid.ixx:17 computes it: a cell exactly on
// an edge (a region's bound, a barrier's x_lo and x_hi) then falls on the same
// side as in the reference.
__device__ __forceinline__ double grid_coord(double lo, long long i, double h) {
#pragma clang fp contract(off)
    return lo + static_cast<double>(i) * h;
}

}  // namespace tdse
