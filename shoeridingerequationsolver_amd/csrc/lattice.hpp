// lattice.hpp -- host view of the Peierls lattice sweeps (2D finite
// differences with link phases, open boundaries).
#pragma once

#include <hip/hip_runtime.h>

namespace tdse {

constexpr int kLatMaxNx = 2048;  // the opening row sweep holds a whole row in LDS (FP64)

struct LatticeArgs {
    void* psi;            // cx<T>, x fastest (the energy reads it; the sweeps take src, dst)
    const void* phi;      // T: -e0 dt / 2 per site, e0 = V + 2 tx + 2 ty
    const void* mask;     // T, stored form (FP32: 1 - m), or null
    const void* v;        // T: V (energy)
    int nx, ny;
    double cx, mxr, mxi;      // x bonds (dt / 2): c and the mixing coefficient
    double cy, myr, myi;      // y bonds of even rows (dt / 2)
    double cy2, my2r, my2i;   // y bonds of odd rows (dt)
    int relax;                // 1: onsite weights exp(f phi), else phases exp(i f phi)
    int link_mode;            // x links: 0 none, 1 uniform field, 2 solenoid
    double link_a;            // uniform: -B hx (angle per unit y), |link_a y| < 1e5
    double ylo, hy;           // y of row j: ylo + j hy
    int link_col, link_j0, link_j1;  // solenoid: column and rows [j0, j1) of the cut
    double link_c, link_s;    // solenoid: the cut's link exp(-i phi) (cut up) or exp(+i phi)
    double tx, ty;            // hopping (energy)
};

// The opening half step of a call (onsite factor, then both x groups), one
// row per block; src and dst may be the same array.
hipError_t launch_lattice_rows(bool fp64, const LatticeArgs& a, const void* src, void* dst,
                               hipStream_t stream);
// `units` (<= kLatStepsPerLaunch) whole steps from one step boundary to the
// next, src -> dst (distinct arrays: tiles read each other's halos): per step
// the y groups, then the x groups around the onsite factor and mask; the last
// step ends with last_stage (1: the next step's x groups follow, 2: the
// call's end).
constexpr int kLatStepsPerLaunch = 2;
hipError_t launch_lattice_steps(bool fp64, const LatticeArgs& a, const void* src, void* dst,
                                int units, int last_stage, hipStream_t stream);
// <H> sums (e, den) into partials[3 b + q] per block, then out[0..2]
hipError_t launch_lattice_energy(bool fp64, const LatticeArgs& a, double* partials, double* out,
                                 hipStream_t stream);

}  // namespace tdse
