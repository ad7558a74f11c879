// device_build.hpp -- host view of the device input builders.
#pragma once

#include <hip/hip_runtime.h>

#include "../../include/tdse.h"

namespace tdse {

struct BuildArgs {
    void* phi;
    void* v;
    void* mask;   // null: no mask
    void* psi;
    long long total;
    int n[3];
    double lo[3], hi[3], h[3];
    double dt;
    int imaginary;  // 1: phi = -V dt / 2 unreduced (imaginary time weights)
    int tan_half;   // 1: phi stored as tan(phi / 2) (FP32 real-time kick sweeps)
    double bfield;  // uniform field: V += B^2/8 |r_perp|^2 (0: none)
    double onsite_shift;  // added to V in the kick phase only (lattice: 2 tx + 2 ty)
    int b_axis;
    tdse_recipe r;
};

hipError_t launch_build(bool fp64, const BuildArgs& a, hipStream_t stream);

}  // namespace tdse
