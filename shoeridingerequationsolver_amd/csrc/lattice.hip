// lattice.hip -- the Peierls lattice propagator (core/src/lattice2d.ixx).
//
// The reference step is  P, X0, X1, Y0, Y1', Y0, X1, X0, P  (P the onsite
// half factor, Xp / Yp the bond groups of parity p, each pair of sites an
// exact 2x2 rotation; Y1' over a full dt).  Consecutive steps meet in
// X1, X0, P, P, X0, X1.  A call opens with one row sweep (P, X0, X1: a row
// in LDS); after that every step is the same unit, Y0 Y1' Y0 X1 X0 P P X0 X1
// (the last step stops after its P), and whole units run on 2D tiles, two
// steps per launch.  Bond rotations and onsite factors are formed in FP64
// for both state types.
#include <hip/hip_runtime.h>

#include "fft_engine.hpp"
#include "lattice.hpp"
#include "observe.hpp"
#include "sweeps.hpp"

namespace tdse {
namespace {

constexpr int kBlock = 256;

__device__ __forceinline__ cx<double> cmuld(cx<double> a, cx<double> b) {
    return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}

// The 2x2 rotation of one bond with link u:
//     pa <- c pa + (mix u) pb,  pb <- (mix conj(u)) pa + c pb.
// mix u and mix conj(u) are the same for every bond of a group that shares a
// link, so they are formed once per group, not per bond.
struct BondCoef {
    double c;
    cx<double> mu, mv;  // mix u, mix conj(u)
};

__device__ __forceinline__ BondCoef bond_coef(double c, cx<double> mix, cx<double> u) {
    return {c, cmuld(mix, u), cmuld(mix, cx<double>{u.x, -u.y})};
}

__device__ __forceinline__ void bond(cx<double>& pa, cx<double>& pb, const BondCoef& k) {
    const cx<double> a0 = pa, b0 = pb;
    const cx<double> t = cmuld(k.mu, b0);
    const cx<double> t2 = cmuld(k.mv, a0);
    pa = {k.c * a0.x + t.x, k.c * a0.y + t.y};
    pb = {t2.x + k.c * b0.x, t2.y + k.c * b0.y};
}

// x link of bond (i, j) -> (i + 1, j)
__device__ __forceinline__ cx<double> link_x(const LatticeArgs& a, int i, int j, cx<double> urow) {
    if (a.link_mode == 1) return urow;
    if (a.link_mode == 2 && i == a.link_col && j >= a.link_j0 && j < a.link_j1) return urow;
    return {1.0, 0.0};
}

// x link of row j: exp(-i B hx y_j) for a uniform field (the host keeps
// |B hx y| below 1e5, inside the short reduction of sincos_small), the
// host-formed exp(+-i phi) on a solenoid's cut
__device__ __forceinline__ cx<double> row_link(const LatticeArgs& a, int j) {
    if (a.link_mode == 1) {
        double s, c;
        sincos_small(a.link_a * (a.ylo + a.hy * j), &s, &c);
        return {c, s};
    }
    if (a.link_mode == 2) return {a.link_c, a.link_s};
    return {1.0, 0.0};
}

// m exp(i f phi) (real time) or m exp(f phi) (imaginary time) of one site
template <typename T>
__device__ __forceinline__ cx<double> onsite_factor(const LatticeArgs& a, T phi, double m,
                                                    double f) {
    const double p = f * static_cast<double>(phi);
    if (a.relax) return {m * exp(p), 0.0};
    double s, c;
    sincos_small(p, &s, &c);
    return {m * c, m * s};
}

// x bonds (i, i + 1), i = parity mod 2, of one row with coefficients k; the
// bond (cut, cut + 1) (a solenoid's branch cut, cut >= 0) takes kcut instead
// and is applied by thread 0 once the others are done with the group's
// disjoint pairs.
__device__ __forceinline__ void x_group(const LatticeArgs& a, cx<double>* row, int parity,
                                        BondCoef k, int cut, BondCoef kcut) {
    for (int b = threadIdx.x;; b += blockDim.x) {
        const int i = 2 * b + parity;
        if (i + 1 >= a.nx) break;
        if (i != cut) bond(row[i], row[i + 1], k);
    }
    if (threadIdx.x == 0 && cut >= 0 && (cut & 1) == parity && cut + 1 < a.nx)
        bond(row[cut], row[cut + 1], kcut);
    __syncthreads();
}

// The opening half step of a call, P X0 X1, one row per block, src -> dst
// (the same array or not).  Thread t owns sites t + u kBlock, u < PER, of a
// row (PER kBlock >= nx); every load of a row is issued before the first use.
template <typename T, int PER>
__global__ void __launch_bounds__(kBlock) lattice_rows(LatticeArgs a, const void* src_,
                                                       void* dst_) {
    __shared__ cx<double> row[kLatMaxNx];
    const cx<T>* src = static_cast<const cx<T>*>(src_);
    cx<T>* dst = static_cast<cx<T>*>(dst_);
    const T* phi = static_cast<const T*>(a.phi);
    for (int j = blockIdx.x; j < a.ny; j += gridDim.x) {
        const long base = static_cast<long>(j) * a.nx;
        cx<T> z[PER];
        T ph[PER];
#pragma unroll
        for (int u = 0; u < PER; ++u) {
            const int i = threadIdx.x + u * kBlock;
            if (i < a.nx) {
                z[u] = src[base + i];
                ph[u] = phi[base + i];
            }
        }
        // x links: every bond of the row (uniform field), or one bond of the
        // rows a solenoid's cut crosses; u = 1 elsewhere
        const cx<double> mix = {a.mxr, a.mxi};
        const bool cut_row = a.link_mode == 2 && j >= a.link_j0 && j < a.link_j1;
        const int cut = cut_row ? a.link_col : -1;
        const cx<double> u = row_link(a, j);
        const BondCoef kx = bond_coef(a.cx, mix, a.link_mode == 1 ? u : cx<double>{1.0, 0.0});
        const BondCoef kcut = bond_coef(a.cx, mix, u);
#pragma unroll
        for (int u = 0; u < PER; ++u) {
            const int i = threadIdx.x + u * kBlock;
            if (i < a.nx) {
                const cx<double> w = {static_cast<double>(z[u].x), static_cast<double>(z[u].y)};
                row[i] = cmuld(w, onsite_factor(a, ph[u], 1.0, 1.0));
            }
            __builtin_amdgcn_sched_barrier(0);  // one site's sine and cosine at a time
        }
        __syncthreads();
        x_group(a, row, 0, kx, cut, kcut);
        x_group(a, row, 1, kx, cut, kcut);
        for (int i = threadIdx.x; i < a.nx; i += blockDim.x)
            dst[base + i] = {static_cast<T>(row[i].x), static_cast<T>(row[i].y)};
        __syncthreads();
    }
}

// ---- whole steps on 2D tiles ----------------------------------------------------
//
// One launch advances K steps past a step boundary: per step the three y
// groups, then X1, X0, the merged onsite factor and mask, and (unless it is
// the call's last step) X0, X1 of the next step.  A tile writes kTileW x
// kTileH sites and holds a window with 3K halo rows and 4K halo columns in
// LDS: each bond group leaves at most one more row (y) or column (x) at a
// window edge stale, and stale sites only spread along their own row or
// column.  Windows start on even columns, so a bond's parity in the window is
// its parity in the lattice.  Windows are clipped at the lattice edges, where
// the open boundary has no bond to miss.  Out of place: tiles read each
// other's halos.

constexpr int kTileW = 32, kTileH = 32;

// K steps per launch, BS threads per tile: 1024 when there are few tiles (4
// waves per SIMD of one tile hide a phase's LDS and FP64 latencies), 512 when
// two tiles per CU fit (their LDS and 128 VGPRs per lane allow it)
template <int K, int BS>
struct StepTile {
    static constexpr int HX = 4 * K, HY = 3 * K;  // halo columns, rows
    static constexpr int W = kTileW + 2 * HX, H = kTileH + 2 * HY;
    static constexpr int PER = (W * H + BS - 1) / BS;             // window sites per thread
    static constexpr int YB = (H / 2 * W + BS - 1) / BS;          // y bonds per thread
    static constexpr int XB = (H * (W / 2) + BS - 1) / BS;        // x bonds per thread
};

// y bonds (j, j + 1), j = parity mod 2, of the window rows [0, wyn) (row 0
// is lattice row wy0) over columns [0, wxn)
template <int W, int YB, int BS>
__device__ __forceinline__ void y_group_tile(cx<double> (*w)[W], int wxn, int wy0, int wyn,
                                             int parity, const BondCoef& k) {
    const int first = (wy0 & 1) != parity ? 1 : 0;
    const int nb = (wyn - first) / 2;
    const int tid = opaque_tid();
#pragma unroll
    for (int u = 0; u < YB; ++u) {
        const int q = tid + u * BS;
        const int lc = q % W, lr = first + 2 * (q / W);
        if (q < nb * W && lc < wxn) bond(w[lr][lc], w[lr + 1][lc], k);
        __builtin_amdgcn_sched_barrier(0);  // one bond in registers at a time
    }
    __syncthreads();
}

// x bonds (i, i + 1), i = parity mod 2, of every window row: xk[row] on the
// row's bonds, kc on a solenoid's cut bond
template <int W, int XB, int BS>
__device__ __forceinline__ void x_group_tile(const LatticeArgs& a, cx<double> (*w)[W],
                                             const BondCoef* xk, const BondCoef* kc, int wx0,
                                             int wxn, int wy0, int wyn, int parity) {
    constexpr int NB = W / 2;
    const int tid = opaque_tid();
#pragma unroll
    for (int u = 0; u < XB; ++u) {
        const int q = tid + u * BS;
        const int lr = q / NB, lc = 2 * (q % NB) + parity;
        if (lr < wyn && lc + 1 < wxn) {
            const int j = wy0 + lr;
            const bool cut = a.link_mode == 2 && wx0 + lc == a.link_col && j >= a.link_j0 &&
                             j < a.link_j1;
            bond(w[lr][lc], w[lr][lc + 1], *(cut ? kc : &xk[lr]));
        }
        __builtin_amdgcn_sched_barrier(0);
    }
    __syncthreads();
}

template <typename T, int K, int BS>
__global__ void __launch_bounds__(BS, 4) lattice_steps(LatticeArgs a, const void* src_, void* dst_,
                                                    int units, int last_stage) {
    using S = StepTile<K, BS>;
    constexpr int W = S::W, H = S::H;
    __shared__ cx<double> w[H][W];
    __shared__ T sph[H][W];
    __shared__ T smk[H][W];
    __shared__ BondCoef xk[H];
    __shared__ BondCoef kc, ky[2];  // the cut bond; y bonds of even and odd rows
    const cx<T>* src = static_cast<const cx<T>*>(src_);
    cx<T>* dst = static_cast<cx<T>*>(dst_);
    const T* phi = static_cast<const T*>(a.phi);
    const T* mask = static_cast<const T*>(a.mask);
    const int tid = threadIdx.x;
    const cx<double> mix = {a.mxr, a.mxi};
    if (tid == 0) {
        kc = bond_coef(a.cx, mix, cx<double>{a.link_c, a.link_s});
        ky[0] = bond_coef(a.cy, cx<double>{a.myr, a.myi}, cx<double>{1.0, 0.0});
        ky[1] = bond_coef(a.cy2, cx<double>{a.my2r, a.my2i}, cx<double>{1.0, 0.0});
    }
    const int xt = (a.nx + kTileW - 1) / kTileW, yt = (a.ny + kTileH - 1) / kTileH;
    for (long t = blockIdx.x; t < static_cast<long>(xt) * yt; t += gridDim.x) {
        const int c0 = static_cast<int>(t % xt) * kTileW, r0 = static_cast<int>(t / xt) * kTileH;
        const int wx0 = c0 - S::HX > 0 ? c0 - S::HX : 0;
        const int wy0 = r0 - S::HY > 0 ? r0 - S::HY : 0;
        const int wxn = (c0 + kTileW + S::HX < a.nx ? c0 + kTileW + S::HX : a.nx) - wx0;
        const int wyn = (r0 + kTileH + S::HY < a.ny ? r0 + kTileH + S::HY : a.ny) - wy0;
        // the window in G rounds: every load of a round (state, phases,
        // mask) is issued before its LDS writes; FP64 at 512 threads takes
        // two rounds to stay within 128 VGPRs
        constexpr int G = sizeof(T) == 8 && BS == 512 ? 2 : 1;
        constexpr int PG = (S::PER + G - 1) / G;
        const int ltid = opaque_tid();
#pragma unroll
        for (int g = 0; g < G; ++g) {
            cx<T> z[PG];
            T ph[PG], mk[PG];
#pragma unroll
            for (int v = 0; v < PG; ++v) {
                const int q = ltid + (g * PG + v) * BS, lr = q / W, lc = q % W;
                if (g * PG + v < S::PER && lr < wyn && lc < wxn) {
                    const long f = static_cast<long>(wy0 + lr) * a.nx + wx0 + lc;
                    z[v] = src[f];
                    ph[v] = phi[f];
                    mk[v] = mask ? mask[f] : kMaskOne<T>;
                }
            }
            if (g == 0 && ltid < wyn) {
                const int j = wy0 + ltid;
                xk[ltid] = bond_coef(a.cx, mix,
                                    a.link_mode == 1 ? row_link(a, j) : cx<double>{1.0, 0.0});
            }
#pragma unroll
            for (int v = 0; v < PG; ++v) {
                const int q = ltid + (g * PG + v) * BS, lr = q / W, lc = q % W;
                if (g * PG + v < S::PER && lr < wyn && lc < wxn) {
                    w[lr][lc] = {static_cast<double>(z[v].x), static_cast<double>(z[v].y)};
                    sph[lr][lc] = ph[v];
                    smk[lr][lc] = mk[v];
                }
            }
        }
        __syncthreads();
        for (int s = 0; s < units; ++s) {
            const int stage = s + 1 < units ? 1 : last_stage;
            y_group_tile<W, S::YB, BS>(w, wxn, wy0, wyn, 0, ky[0]);
            y_group_tile<W, S::YB, BS>(w, wxn, wy0, wyn, 1, ky[1]);
            y_group_tile<W, S::YB, BS>(w, wxn, wy0, wyn, 0, ky[0]);
            x_group_tile<W, S::XB, BS>(a, w, xk, &kc, wx0, wxn, wy0, wyn, 1);
            x_group_tile<W, S::XB, BS>(a, w, xk, &kc, wx0, wxn, wy0, wyn, 0);
            // end of one step (and its mask), start of the next: merged
            const double f = stage == 1 ? 2.0 : 1.0;
            const int ptid = opaque_tid();
#pragma unroll
            for (int u = 0; u < S::PER; ++u) {
                const int q = ptid + u * BS, lr = q / W, lc = q % W;
                if (lr < wyn && lc < wxn)
                    w[lr][lc] = cmuld(w[lr][lc], onsite_factor(a, sph[lr][lc],
                                                               mask_value(smk[lr][lc]), f));
                __builtin_amdgcn_sched_barrier(0);
            }
            __syncthreads();
            if (stage == 1) {
                x_group_tile<W, S::XB, BS>(a, w, xk, &kc, wx0, wxn, wy0, wyn, 0);
                x_group_tile<W, S::XB, BS>(a, w, xk, &kc, wx0, wxn, wy0, wyn, 1);
            }
        }
        for (int q = opaque_tid(); q < kTileW * kTileH; q += BS) {
            const int r = r0 + q / kTileW, c = c0 + q % kTileW;
            if (r < a.ny && c < a.nx) {
                const cx<double> v = w[r - wy0][c - wx0];
                dst[static_cast<long>(r) * a.nx + c] = {static_cast<T>(v.x), static_cast<T>(v.y)};
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

hipError_t launch_lattice_rows(bool fp64, const LatticeArgs& a, const void* src, void* dst,
                               hipStream_t stream) {
    if (a.nx < 2 || a.nx > kLatMaxNx) return hipErrorInvalidValue;
    const int blocks = blocks_for(a.ny, 8192);
    const int per = (a.nx + kBlock - 1) / kBlock;
    auto go = [&](auto kernel) {
        hipLaunchKernelGGL(kernel, dim3(blocks), dim3(kBlock), 0, stream, a, src, dst);
    };
    if (fp64) {
        if (per == 1) go(lattice_rows<double, 1>);
        else if (per == 2) go(lattice_rows<double, 2>);
        else if (per <= 4) go(lattice_rows<double, 4>);
        else go(lattice_rows<double, kLatMaxNx / kBlock>);
    } else {
        if (per == 1) go(lattice_rows<float, 1>);
        else if (per == 2) go(lattice_rows<float, 2>);
        else if (per <= 4) go(lattice_rows<float, 4>);
        else go(lattice_rows<float, kLatMaxNx / kBlock>);
    }
    return hipGetLastError();
}

hipError_t launch_lattice_steps(bool fp64, const LatticeArgs& a, const void* src, void* dst,
                                int units, int last_stage, hipStream_t stream) {
    if (units < 1 || units > kLatStepsPerLaunch || (last_stage != 1 && last_stage != 2))
        return hipErrorInvalidValue;
    const long long tiles = static_cast<long long>((a.nx + kTileW - 1) / kTileW) *
                            ((a.ny + kTileH - 1) / kTileH);
    const int blocks = blocks_for(tiles, 8192);
    auto go = [&](auto kernel, int bs) {
        hipLaunchKernelGGL(kernel, dim3(blocks), dim3(bs), 0, stream, a, src, dst, units,
                           last_stage);
    };
    constexpr int KL = kLatStepsPerLaunch;
    if (tiles >= 512) {  // two tiles per CU on 256 CUs
        if (units == 1) fp64 ? go(lattice_steps<double, 1, 512>, 512) : go(lattice_steps<float, 1, 512>, 512);
        else fp64 ? go(lattice_steps<double, KL, 512>, 512) : go(lattice_steps<float, KL, 512>, 512);
    } else {
        if (units == 1) fp64 ? go(lattice_steps<double, 1, 1024>, 1024) : go(lattice_steps<float, 1, 1024>, 1024);
        else fp64 ? go(lattice_steps<double, KL, 1024>, 1024) : go(lattice_steps<float, KL, 1024>, 1024);
    }
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
