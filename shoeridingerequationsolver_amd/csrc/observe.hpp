// observe.hpp -- host view of the reduction kernels.
#pragma once

#include <hip/hip_runtime.h>

#include "fft_engine.hpp"

namespace tdse {

constexpr int kReduceQ = 9;
constexpr int kMaxReduceBlocks = 2048;

struct ReduceArgs {
    const void* psi;   // state (x space) or its forward transform (k space)
    const void* v;     // potential in path precision, or null
    long long total;   // points
    int n[3];          // original dims (x space coordinates)
    double lo[3], h[3];
    // box of the region probability (slot 5): box_lo <= coordinate < box_hi
    // on every axis (infinite bounds leave an axis unrestricted)
    double box_lo[3], box_hi[3];
    int sq[3];         // squeezed dims (k space)
    double dk[3];      // 2 pi / L per squeezed axis
    double mass;
    long long four_step_rows;  // N2 of a four-step 1D grid, else 0
};

// the region probability's box: empty (no region asked for), or every axis
// unrestricted before the caller bounds some
inline void empty_box(ReduceArgs& r) {
    for (int a = 0; a < 3; ++a) {
        r.box_lo[a] = 1.0;
        r.box_hi[a] = 0.0;
    }
}
inline void open_box(ReduceArgs& r) {
    for (int a = 0; a < 3; ++a) {
        r.box_lo[a] = -__builtin_huge_val();
        r.box_hi[a] = __builtin_huge_val();
    }
}

int reduce_blocks(long long total);
// partials: reduce_blocks(total) * kReduceQ doubles; out: kReduceQ doubles (device)
hipError_t launch_reduce(bool fp64, bool kspace, const ReduceArgs& a, double* partials,
                         double* out, hipStream_t stream);
hipError_t launch_scale(bool fp64, void* psi, long long total, double s, hipStream_t stream);
// out[q] = sum over blocks b < nblocks of partials[3 b + q], q < 3, fixed order
// max |psi|^2 over `total` points into out[0] (partials: reduce_blocks(total) doubles)
hipError_t launch_peak(bool fp64, const void* psi, long long total, double* partials, double* out,
                       hipStream_t stream);
hipError_t launch_reduce_final3(const double* partials, int nblocks, double* out,
                                hipStream_t stream);
// Energy variance sums from psi, its kinetic part T psi and V (or null):
// out[0], out[1] = sum |psi|^2, sum Re(conj(psi) H psi); out[kReduceQ] =
// sum |H psi - <H> psi|^2 (2 kReduceQ doubles of device memory)
hipError_t launch_energy_variance(bool fp64, const void* psi, const void* tpsi, const void* v,
                                 long long total, double* partials, double* out,
                                 hipStream_t stream);
// psi *= *s (device scalar)
hipError_t launch_scale_dev(bool fp64, void* psi, long long total, const double* s,
                            hipStream_t stream);
// out[0] + i out[1] = sum conj(a) b (device; partials as launch_reduce)
hipError_t launch_inner(bool fp64, const void* a, const void* b, long long total, double* partials,
                        double* out, hipStream_t stream);
// psi -= (sum[0] + i sum[1]) dv * phi
hipError_t launch_subtract_projection(bool fp64, void* psi, const void* phi, long long total,
                                      const double* sum, double dv, hipStream_t stream);
// psi /= sqrt(sum[0] dv) unless zero; *norm_out = sum[0] dv
hipError_t launch_normalize_dev(bool fp64, void* psi, long long total, const double* sum, double dv,
                                double* norm_out, hipStream_t stream);
// Imaginary time: after a measuring sweep with args `a`, turn its per-block
// partials into the next scale, the norm and the peak (a LaunchFn)
struct SweepArgs;
hipError_t launch_relax_finish(const SweepArgs& a, int max_blocks, hipStream_t stream);
// Streaming copy kernel (16-byte lanes), the bandwidth reference
hipError_t launch_copy(void* dst, const void* src, size_t bytes, hipStream_t stream);
// FP64 elementwise passes of the unfused baseline
hipError_t launch_table_mul(void* psi, const void* table, long long total, hipStream_t stream);
hipError_t launch_conj_scale(void* psi, long long total, double s, hipStream_t stream);
hipError_t launch_mask_mul(void* psi, const double* mask, long long total, hipStream_t stream);

}  // namespace tdse
