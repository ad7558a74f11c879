// lattice.hip -- the Peierls lattice propagator (core/src/lattice2d.ixx) as
// two sweeps per step.
//
// The reference step is  P, X0, X1, Y0, Y1', Y0, X1, X0, P  (P the onsite
// half factor, Xp / Yp the bond groups of parity p, each pair of sites an
// exact 2x2 rotation; Y1' over a full dt).  Consecutive steps meet in
// X1, X0, P, P, X0, X1, so a step boundary is ONE row sweep (a row in LDS,
// both x groups, the merged onsite factor and the caller's mask) and the Y
// groups are ONE column sweep (a tile of columns with a 3-row halo: a final
// row value depends on rows j-3 .. j+3 through the three groups).  Tiles
// read each other's halo rows, so the column sweep writes a second array
// and the next row sweep reads it back.  Bond rotations and onsite factors
// are formed in FP64 for both state types.
#include <hip/hip_runtime.h>

#include "fft_engine.hpp"
#include "lattice.hpp"
#include "observe.hpp"
#include "sweeps.hpp"

namespace tdse {
namespace {

constexpr int kBlock = 256;
constexpr int kCols = 32;   // columns per column tile
constexpr int kRows = 58;   // rows written per column tile
constexpr int kHalo = 3;
constexpr int kWin = kRows + 2 * kHalo;

__device__ __forceinline__ cx<double> cmuld(cx<double> a, cx<double> b) {
    return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}

// pa <- c pa + (mix u) pb,  pb <- (mix conj(u)) pa + c pb
__device__ __forceinline__ void bond(cx<double>& pa, cx<double>& pb, double c, cx<double> mix,
                                     cx<double> u) {
    const cx<double> a0 = pa, b0 = pb;
    const cx<double> t = cmuld(cmuld(mix, u), b0);
    const cx<double> t2 = cmuld(cmuld(mix, cx<double>{u.x, -u.y}), a0);
    pa = {c * a0.x + t.x, c * a0.y + t.y};
    pb = {t2.x + c * b0.x, t2.y + c * b0.y};
}

// x link of bond (i, j) -> (i + 1, j)
__device__ __forceinline__ cx<double> link_x(const LatticeArgs& a, int i, int j, cx<double> urow) {
    if (a.link_mode == 1) return urow;
    if (a.link_mode == 2 && i == a.link_col && j >= a.link_j0 && j < a.link_j1) return urow;
    return {1.0, 0.0};
}

__device__ __forceinline__ cx<double> row_link(const LatticeArgs& a, int j) {
    double s, c;
    if (a.link_mode == 1) {
        sincos(a.link_a * (a.ylo + a.hy * j), &s, &c);
        return {c, s};
    }
    if (a.link_mode == 2) {
        sincos(a.link_a, &s, &c);
        return {c, a.link_sign * s};
    }
    return {1.0, 0.0};
}

template <typename T>
__device__ __forceinline__ void onsite(const LatticeArgs& a, cx<double>* row, long base, double f,
                                       bool with_mask) {
    const T* phi = static_cast<const T*>(a.phi);
    const T* mask = static_cast<const T*>(a.mask);
    for (int i = threadIdx.x; i < a.nx; i += blockDim.x) {
        const double p = f * static_cast<double>(phi[base + i]);
        double m = 1.0;
        if (with_mask && mask) m = mask_value(mask[base + i]);
        cx<double> fac;
        if (a.relax) {
            fac = {m * exp(p), 0.0};
        } else {
            double s, c;
            sincos_small(p, &s, &c);
            fac = {m * c, m * s};
        }
        row[i] = cmuld(row[i], fac);
    }
}

__device__ __forceinline__ void x_group(const LatticeArgs& a, cx<double>* row, int j, int parity,
                                        cx<double> urow) {
    const cx<double> mix = {a.mxr, a.mxi};
    for (int b = threadIdx.x;; b += blockDim.x) {
        const int i = 2 * b + parity;
        if (i + 1 >= a.nx) break;
        bond(row[i], row[i + 1], a.cx, mix, link_x(a, i, j, urow));
    }
    __syncthreads();
}

template <typename T>
__global__ void __launch_bounds__(kBlock) lattice_rows(LatticeArgs a) {
    __shared__ cx<double> row[kLatMaxNx];
    cx<T>* psi = static_cast<cx<T>*>(a.psi);
    const cx<T>* src = static_cast<const cx<T>*>(a.other ? a.other : a.psi);
    for (int j = blockIdx.x; j < a.ny; j += gridDim.x) {
        const long base = static_cast<long>(j) * a.nx;
        const cx<double> urow = row_link(a, j);
        for (int i = threadIdx.x; i < a.nx; i += blockDim.x) {
            const cx<T> z = src[base + i];
            row[i] = {static_cast<double>(z.x), static_cast<double>(z.y)};
        }
        __syncthreads();
        if (a.stage == 0) {
            onsite<T>(a, row, base, 1.0, false);
            __syncthreads();
            x_group(a, row, j, 0, urow);
            x_group(a, row, j, 1, urow);
        } else {
            x_group(a, row, j, 1, urow);
            x_group(a, row, j, 0, urow);
            // end of one step (and its mask), start of the next: merged
            onsite<T>(a, row, base, a.stage == 1 ? 2.0 : 1.0, true);
            __syncthreads();
            if (a.stage == 1) {
                x_group(a, row, j, 0, urow);
                x_group(a, row, j, 1, urow);
            }
        }
        for (int i = threadIdx.x; i < a.nx; i += blockDim.x)
            psi[base + i] = {static_cast<T>(row[i].x), static_cast<T>(row[i].y)};
        __syncthreads();
    }
}

// y bonds (j, j + 1) of rows j = parity mod 2 inside the window [wlo, whi)
__device__ __forceinline__ void y_group(cx<double> (*tile)[kCols + 1], int wlo, int whi, int parity,
                                        double c, cx<double> mix) {
    const int first = wlo + ((wlo & 1) != parity ? 1 : 0);
    const int nb = first + 1 < whi ? (whi - 1 - first + 1) / 2 : 0;  // bonds j = first, first+2, ..
    for (int q = threadIdx.x; q < nb * kCols; q += blockDim.x) {
        const int col = q % kCols, jb = first + 2 * (q / kCols);
        if (jb + 1 >= whi) continue;
        bond(tile[jb - wlo][col], tile[jb + 1 - wlo][col], c, mix, cx<double>{1.0, 0.0});
    }
    __syncthreads();
}

template <typename T>
__global__ void __launch_bounds__(kBlock) lattice_cols(LatticeArgs a) {
    __shared__ cx<double> tile[kWin][kCols + 1];
    const cx<T>* psi = static_cast<const cx<T>*>(a.psi);
    cx<T>* out = static_cast<cx<T>*>(a.other);
    const int xt = (a.nx + kCols - 1) / kCols, yt = (a.ny + kRows - 1) / kRows;
    for (long t = blockIdx.x; t < static_cast<long>(xt) * yt; t += gridDim.x) {
        const int c0 = static_cast<int>(t % xt) * kCols, r0 = static_cast<int>(t / xt) * kRows;
        const int wlo = r0 - kHalo > 0 ? r0 - kHalo : 0;
        const int whi = r0 + kRows + kHalo < a.ny ? r0 + kRows + kHalo : a.ny;
        for (int q = threadIdx.x; q < (whi - wlo) * kCols; q += blockDim.x) {
            const int col = q % kCols, r = wlo + q / kCols;
            cx<double> z = {0.0, 0.0};
            if (c0 + col < a.nx) {
                const cx<T> v = psi[static_cast<long>(r) * a.nx + c0 + col];
                z = {static_cast<double>(v.x), static_cast<double>(v.y)};
            }
            tile[r - wlo][col] = z;
        }
        __syncthreads();
        y_group(tile, wlo, whi, 0, a.cy, cx<double>{a.myr, a.myi});
        y_group(tile, wlo, whi, 1, a.cy2, cx<double>{a.my2r, a.my2i});
        y_group(tile, wlo, whi, 0, a.cy, cx<double>{a.myr, a.myi});
        const int rend = r0 + kRows < a.ny ? r0 + kRows : a.ny;
        for (int q = threadIdx.x; q < (rend - r0) * kCols; q += blockDim.x) {
            const int col = q % kCols, r = r0 + q / kCols;
            if (c0 + col < a.nx) {
                const cx<double> z = tile[r - wlo][col];
                out[static_cast<long>(r) * a.nx + c0 + col] = {static_cast<T>(z.x),
                                                               static_cast<T>(z.y)};
            }
        }
        __syncthreads();
    }
}

// core/src/lattice2d.ixx:251-274: sum of (V + 2 tx + 2 ty)|z|^2 and the
// bond terms -2 t Re(conj(z) u z'), and sum |z|^2, per block
template <typename T>
__global__ void __launch_bounds__(kBlock) lattice_energy(LatticeArgs a, double* partials) {
    const cx<T>* psi = static_cast<const cx<T>*>(a.psi);
    const T* v = static_cast<const T*>(a.v);
    double e = 0.0, den = 0.0;
    const long total = static_cast<long>(a.nx) * a.ny;
    for (long f = blockIdx.x * static_cast<long>(kBlock) + threadIdx.x; f < total;
         f += static_cast<long>(gridDim.x) * kBlock) {
        const int i = static_cast<int>(f % a.nx), j = static_cast<int>(f / a.nx);
        const cx<double> z = {static_cast<double>(psi[f].x), static_cast<double>(psi[f].y)};
        const cx<double> zc = {z.x, -z.y};
        const double w = z.x * z.x + z.y * z.y;
        den += w;
        e += (static_cast<double>(v[f]) + 2.0 * a.tx + 2.0 * a.ty) * w;
        if (i + 1 < a.nx) {
            const cx<double> zn = {static_cast<double>(psi[f + 1].x), static_cast<double>(psi[f + 1].y)};
            e += -2.0 * a.tx * cmuld(cmuld(zc, link_x(a, i, j, row_link(a, j))), zn).x;
        }
        if (j + 1 < a.ny) {
            const cx<double> zn = {static_cast<double>(psi[f + a.nx].x),
                                   static_cast<double>(psi[f + a.nx].y)};
            e += -2.0 * a.ty * cmuld(zc, zn).x;
        }
    }
    __shared__ double red[2 * (kBlock / 64)];
    for (int off = 32; off > 0; off >>= 1) {
        e += __shfl_xor(e, off, 64);
        den += __shfl_xor(den, off, 64);
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) {
        red[2 * wave] = e;
        red[2 * wave + 1] = den;
    }
    __syncthreads();
    if (threadIdx.x < 3) {
        double s = 0.0;
        if (threadIdx.x < 2)
            for (int w = 0; w < kBlock / 64; ++w) s += red[2 * w + threadIdx.x];
        partials[3 * blockIdx.x + threadIdx.x] = s;
    }
}

int blocks_for(long long work, int cap) {
    long long b = work < 1 ? 1 : work;
    return static_cast<int>(b < cap ? b : cap);
}

}  // namespace

hipError_t launch_lattice_rows(bool fp64, const LatticeArgs& a, hipStream_t stream) {
    const int blocks = blocks_for(a.ny, 8192);
    if (fp64)
        hipLaunchKernelGGL(lattice_rows<double>, dim3(blocks), dim3(kBlock), 0, stream, a);
    else
        hipLaunchKernelGGL(lattice_rows<float>, dim3(blocks), dim3(kBlock), 0, stream, a);
    return hipGetLastError();
}

hipError_t launch_lattice_cols(bool fp64, const LatticeArgs& a, hipStream_t stream) {
    const long long tiles = static_cast<long long>((a.nx + kCols - 1) / kCols) *
                            ((a.ny + kRows - 1) / kRows);
    const int blocks = blocks_for(tiles, 8192);
    if (fp64)
        hipLaunchKernelGGL(lattice_cols<double>, dim3(blocks), dim3(kBlock), 0, stream, a);
    else
        hipLaunchKernelGGL(lattice_cols<float>, dim3(blocks), dim3(kBlock), 0, stream, a);
    return hipGetLastError();
}

hipError_t launch_lattice_energy(bool fp64, const LatticeArgs& a, double* partials, double* out,
                                 hipStream_t stream) {
    const long long total = static_cast<long long>(a.nx) * a.ny;
    const int blocks = blocks_for((total + kBlock * 8 - 1) / (kBlock * 8), kMaxReduceBlocks);
    if (fp64)
        hipLaunchKernelGGL(lattice_energy<double>, dim3(blocks), dim3(kBlock), 0, stream, a, partials);
    else
        hipLaunchKernelGGL(lattice_energy<float>, dim3(blocks), dim3(kBlock), 0, stream, a, partials);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    return launch_reduce_final3(partials, blocks, out, stream);
}

}  // namespace tdse
