// lattice.hpp -- host view of the Peierls lattice sweeps (2D finite
// differences with link phases, open boundaries).
#pragma once

#include <hip/hip_runtime.h>

namespace tdse {

constexpr int kLatMaxNx = 2048;  // a whole row sits in LDS (FP64)

struct LatticeArgs {
    void* psi;            // cx<T>, x fastest
    // The column sweep reads psi and writes `other` (its halo rows are read
    // by neighbouring tiles, so it cannot work in place); the row sweep then
    // reads `other` (null: psi) and writes psi.
    void* other;
    const void* phi;      // T: -e0 dt / 2 per site, e0 = V + 2 tx + 2 ty
    const void* mask;     // T, stored form (FP32: 1 - m), or null
    const void* v;        // T: V (energy)
    int nx, ny;
    double cx, mxr, mxi;      // x bonds (dt / 2): c and the mixing coefficient
    double cy, myr, myi;      // y bonds of even rows (dt / 2)
    double cy2, my2r, my2i;   // y bonds of odd rows (dt)
    int relax;                // 1: onsite weights exp(f phi), else phases exp(i f phi)
    int link_mode;            // x links: 0 none, 1 uniform field, 2 solenoid
    double link_a;            // uniform: -B hx (angle per unit y); solenoid: the flux phi
    double ylo, hy;           // y of row j: ylo + j hy
    int link_col, link_j0, link_j1;  // solenoid: column and rows [j0, j1) of the cut
    double link_sign;         // solenoid: -1 (cut up, exp(-i phi)) or +1 (cut down)
    int stage;                // row sweeps: 0 first, 1 between steps, 2 last
    double tx, ty;            // hopping (energy)
};

// One row sweep (x bonds and the onsite factor of `stage`) or one column
// sweep (the three y-bond groups of a step).
hipError_t launch_lattice_rows(bool fp64, const LatticeArgs& a, hipStream_t stream);
hipError_t launch_lattice_cols(bool fp64, const LatticeArgs& a, hipStream_t stream);
// <H> sums (e, den) into partials[3 b + q] per block, then out[0..2]
hipError_t launch_lattice_energy(bool fp64, const LatticeArgs& a, double* partials, double* out,
                                 hipStream_t stream);

}  // namespace tdse
