// device_build.hip -- build a plan's inputs in HBM from a recipe.
//
// The same formulas as the host builders (core/src/potential.ixx,
// core/src/wavepacket.ixx) evaluated per cell in FP64 on the GPU, so a served
// request needs no host-side array and no PCIe upload.  Results agree with
// the host builders to a few ulps (device exp/pow/sin), not bit for bit; the
// host builders remain the bit-exact path.
#include <hip/hip_runtime.h>

#include "../../include/tdse.h"
#include "device_build.hpp"
#include "fft_engine.hpp"
#include "grid.hpp"

namespace tdse {
namespace {

constexpr int kBlock = 256;
constexpr double kPi = 3.14159265358979323846;

__device__ __forceinline__ double reduce_2pi_dev(double th) {
    constexpr double kTwoPiHi = 6.28318530717958623200e+00;
    constexpr double kTwoPiLo = 2.44929359829470635445e-16;
    const double n = rint(th * 0.15915494309189533577);
    return (th - n * kTwoPiHi) - n * kTwoPiLo;
}

__device__ __forceinline__ double ramp(double x, double lo, double hi, int n, double width) {
    if (n == 1) return 1.0;
    const double d_lo = x - lo, d_hi = hi - x;
    const double d = d_lo < d_hi ? d_lo : d_hi;
    if (d >= width) return 1.0;
    const double s = sin(0.5 * kPi * (d / width));
    return s * s;
}

__device__ __forceinline__ double envelope(double u, double u0, double s) {
    return pow(2.0 * kPi * s * s, -0.25) * exp(-(u - u0) * (u - u0) / (4.0 * s * s));
}

template <typename T>
__global__ void __launch_bounds__(kBlock)
build_cells(BuildArgs a) {
    T* phi = static_cast<T*>(a.phi);
    T* v = static_cast<T*>(a.v);
    T* mask = static_cast<T*>(a.mask);
    cx<T>* psi = static_cast<cx<T>*>(a.psi);
    const tdse_recipe& r = a.r;
    const long long stride = static_cast<long long>(gridDim.x) * kBlock;
    for (long long f = blockIdx.x * static_cast<long long>(kBlock) + threadIdx.x; f < a.total;
         f += stride) {
        const int i = static_cast<int>(f % a.n[0]);
        const long long q = f / a.n[0];
        const int j = static_cast<int>(q % a.n[1]);
        const int k = static_cast<int>(q / a.n[1]);
        const double x = grid_coord(a.lo[0], i, a.h[0]), y = grid_coord(a.lo[1], j, a.h[1]),
                     z = grid_coord(a.lo[2], k, a.h[2]);
        const double dx = x - r.center[0], dy = y - r.center[1], dz = z - r.center[2];
        double pot = 0.0;
        switch (r.potential) {
            case TDSE_POT_HARMONIC:
                pot = 0.5 * r.strength * r.strength * (dx * dx + dy * dy + dz * dz);
                break;
            case TDSE_POT_REGULARIZED_COULOMB: {
                const double rr = sqrt(dx * dx + dy * dy + dz * dz);
                pot = (rr < 1e-6 * a.h[0]) ? -r.strength * 2.3800774 / a.h[0] : -r.strength / rr;
                break;
            }
            case TDSE_POT_COULOMB_FLOOR: {
                const double rr = sqrt(dx * dx + dy * dy + dz * dz);
                const double rmin = 0.5 * a.h[0];
                pot = -r.strength / (rr > rmin ? rr : rmin);
                break;
            }
            case TDSE_POT_BARRIER:
                pot = (x >= r.x_lo && x < r.x_hi) ? r.strength : 0.0;
                break;
            case TDSE_POT_SOFT_COULOMB:
                pot = -r.strength / sqrt(dx * dx + dy * dy + dz * dz + r.softening * r.softening);
                break;
            default:
                pot = 0.0;
        }
        if (a.bfield != 0.0) {
            const double c[3] = {x, y, z};
            double perp2 = 0.0;
            for (int q = 0; q < 3; ++q)
                if (q != a.b_axis) perp2 += c[q] * c[q];
            pot += a.bfield * a.bfield / 8.0 * perp2;
        }
        v[f] = static_cast<T>(pot);
        const double e0 = pot + a.onsite_shift;
        double kick = a.imaginary ? -0.5 * e0 * a.dt : reduce_2pi_dev(-0.5 * e0 * a.dt);
        if (a.tan_half) kick = tan(0.5 * kick);
        phi[f] = static_cast<T>(kick);
        if (mask) {
            const double m = ramp(x, a.lo[0], a.hi[0], a.n[0], r.mask_width) *
                             ramp(y, a.lo[1], a.hi[1], a.n[1], r.mask_width) *
                             ramp(z, a.lo[2], a.hi[2], a.n[2], r.mask_width);
            // FP32 plans store 1 - m (full precision just below m = 1)
            mask[f] = static_cast<T>(sizeof(T) == 4 ? 1.0 - m : m);
        }
        const double env = envelope(x, r.r0[0], r.sigma[0]) * envelope(y, r.r0[1], r.sigma[1]) *
                           envelope(z, r.r0[2], r.sigma[2]);
        const double ph = r.k0[0] * x + r.k0[1] * y + r.k0[2] * z;
        double s, c;
        sincos(ph, &s, &c);
        psi[f] = {static_cast<T>(env * c), static_cast<T>(env * s)};
    }
}

}  // namespace

hipError_t launch_build(bool fp64, const BuildArgs& a, hipStream_t stream) {
    long long b = (a.total + kBlock * 4 - 1) / (kBlock * 4);
    if (b < 1) b = 1;
    if (b > 16384) b = 16384;
    if (fp64)
        hipLaunchKernelGGL(build_cells<double>, dim3(b), dim3(kBlock), 0, stream, a);
    else
        hipLaunchKernelGGL(build_cells<float>, dim3(b), dim3(kBlock), 0, stream, a);
    return hipGetLastError();
}

}  // namespace tdse
