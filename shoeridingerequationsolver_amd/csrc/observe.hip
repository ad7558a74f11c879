// observe.hip -- deterministic FP64 reductions over the resident state.
//
// Every block walks a fixed, grid-stride set of points in a fixed order,
// accumulates in FP64, reduces inside each wave64 with a fixed xor-shuffle
// tree and across waves through LDS in wave order, and writes one partial per
// block.  A single-block pass then sums the partials in index order.  For a
// given shape the result is bitwise reproducible from run to run (no atomics),
// which the reference's deterministic pool guarantees on the CPU
// (core/src/parallel.ixx:14-16, 211-230).
#include <hip/hip_runtime.h>

#include "grid.hpp"
#include "observe.hpp"
#include "sweeps.hpp"

namespace tdse {
namespace {

constexpr int kBlock = 256;
constexpr int kWaves = kBlock / 64;

template <int Q>
__device__ __forceinline__ void block_reduce_store(double (&acc)[Q], double* out) {
    __shared__ double part[kWaves][Q];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        double v = acc[q];
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
        acc[q] = v;
    }
    if (lane == 0) {
#pragma unroll
        for (int q = 0; q < Q; ++q) part[wave][q] = acc[q];
    }
    __syncthreads();
    if (threadIdx.x < Q) {
        double s = 0.0;
#pragma unroll
        for (int w = 0; w < kWaves; ++w) s += part[w][threadIdx.x];
        out[threadIdx.x] = s;
    }
}

__device__ __forceinline__ int shift_k(int j, int n) { return (2 * j < n) ? j : j - n; }

// x-space sums: |psi|^2, x|psi|^2, y|psi|^2, z|psi|^2, V|psi|^2, box |psi|^2,
// x^2|psi|^2, y^2|psi|^2, z^2|psi|^2
template <typename T>
__global__ void __launch_bounds__(kBlock)
reduce_x(const ReduceArgs a, double* partials) {
    const cx<T>* psi = static_cast<const cx<T>*>(a.psi);
    const T* v = static_cast<const T*>(a.v);
    double acc[kReduceQ] = {};
    const long long stride = static_cast<long long>(gridDim.x) * kBlock;
    for (long long f = blockIdx.x * static_cast<long long>(kBlock) + threadIdx.x; f < a.total;
         f += stride) {
        const cx<T> z = psi[f];
        const double w = static_cast<double>(z.x) * z.x + static_cast<double>(z.y) * z.y;
        const int ix = static_cast<int>(f % a.n[0]);
        const long long r = f / a.n[0];
        const int iy = static_cast<int>(r % a.n[1]);
        const int iz = static_cast<int>(r / a.n[1]);
        const double cxv = grid_coord(a.lo[0], ix, a.h[0]);
        const double cyv = grid_coord(a.lo[1], iy, a.h[1]);
        const double czv = grid_coord(a.lo[2], iz, a.h[2]);
        acc[0] += w;
        acc[1] += cxv * w;
        acc[2] += cyv * w;
        acc[3] += czv * w;
        if (v) acc[4] += static_cast<double>(v[f]) * w;
        if (cxv >= a.box_lo[0] && cxv < a.box_hi[0] && cyv >= a.box_lo[1] && cyv < a.box_hi[1] &&
            czv >= a.box_lo[2] && czv < a.box_hi[2])
            acc[5] += w;
        acc[6] += cxv * cxv * w;
        acc[7] += cyv * cyv * w;
        acc[8] += czv * czv * w;
    }
    block_reduce_store<kReduceQ>(acc, partials + blockIdx.x * kReduceQ);
}

// k-space sums of the forward transform: |phi|^2, 0.5 |k|^2 / m |phi|^2 and
// the first moments k_a |phi|^2 (slots 2-4).
// Natural order on the squeezed axes, or the four-step order of a 1D grid
// (address k1 + N1*k2 holds bin k2 + N2*k1).
template <typename T>
__global__ void __launch_bounds__(kBlock)
reduce_k(const ReduceArgs a, double* partials) {
    const cx<T>* phi = static_cast<const cx<T>*>(a.psi);
    double acc[kReduceQ] = {};
    const long long stride = static_cast<long long>(gridDim.x) * kBlock;
    for (long long f = blockIdx.x * static_cast<long long>(kBlock) + threadIdx.x; f < a.total;
         f += stride) {
        const cx<T> z = phi[f];
        const double w = static_cast<double>(z.x) * z.x + static_cast<double>(z.y) * z.y;
        double k2sum;
        if (a.four_step_rows > 0) {
            const long long k1 = f % a.sq[0];
            const long long k2 = f / a.sq[0];
            const int j = static_cast<int>(k2 + a.four_step_rows * k1);
            const double k = a.dk[0] * shift_k(j, static_cast<int>(a.total));
            k2sum = k * k;
            acc[2] += k * w;
        } else {
            const int i0 = static_cast<int>(f % a.sq[0]);
            const long long r = f / a.sq[0];
            const int i1 = static_cast<int>(r % a.sq[1]);
            const int i2 = static_cast<int>(r / a.sq[1]);
            const double k0 = a.dk[0] * shift_k(i0, a.sq[0]);
            const double k1 = a.dk[1] * shift_k(i1, a.sq[1]);
            const double k2 = a.dk[2] * shift_k(i2, a.sq[2]);
            k2sum = k0 * k0 + k1 * k1 + k2 * k2;
            acc[2] += k0 * w;
            acc[3] += k1 * w;
            acc[4] += k2 * w;
        }
        acc[0] += w;
        acc[1] += 0.5 * k2sum / a.mass * w;
    }
    block_reduce_store<kReduceQ>(acc, partials + blockIdx.x * kReduceQ);
}

__global__ void __launch_bounds__(kBlock)
reduce_final(const double* partials, int nparts, double* out) {
    double acc[kReduceQ] = {};
    for (int p = threadIdx.x; p < nparts; p += kBlock) {
#pragma unroll
        for (int q = 0; q < kReduceQ; ++q) acc[q] += partials[p * kReduceQ + q];
    }
    block_reduce_store<kReduceQ>(acc, out);
}

template <typename T>
__global__ void __launch_bounds__(kBlock)
scale_state(void* psi_v, long long total, double s) {
    cx<T>* psi = static_cast<cx<T>*>(psi_v);
    const T ss = static_cast<T>(s);
    const long long stride = static_cast<long long>(gridDim.x) * kBlock;
    for (long long f = blockIdx.x * static_cast<long long>(kBlock) + threadIdx.x; f < total;
         f += stride) {
        cx<T> z = psi[f];
        psi[f] = {z.x * ss, z.y * ss};
    }
}

// max |psi|^2 per block, then over the blocks (a max: the same bits in any
// order).
template <typename T>
__global__ void __launch_bounds__(kBlock)
reduce_peak(const void* psi_v, long long total, double* partials) {
    const cx<T>* psi = static_cast<const cx<T>*>(psi_v);
    double m = 0.0;
    const long long stride = static_cast<long long>(gridDim.x) * kBlock;
    for (long long f = blockIdx.x * static_cast<long long>(kBlock) + threadIdx.x; f < total;
         f += stride) {
        const cx<T> z = psi[f];
        const double w = static_cast<double>(z.x) * z.x + static_cast<double>(z.y) * z.y;
        m = w > m ? w : m;
    }
    for (int off = 32; off > 0; off >>= 1) {
        const double o = __shfl_xor(m, off, 64);
        m = o > m ? o : m;
    }
    __shared__ double part[kWaves];
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < kWaves; ++w) m = part[w] > m ? part[w] : m;
        partials[blockIdx.x] = part[0] > m ? part[0] : m;
    }
}

__global__ void __launch_bounds__(kBlock)
reduce_peak_final(const double* partials, int nparts, double* out) {
    double m = 0.0;
    for (int p = threadIdx.x; p < nparts; p += kBlock) m = partials[p] > m ? partials[p] : m;
    for (int off = 32; off > 0; off >>= 1) {
        const double o = __shfl_xor(m, off, 64);
        m = o > m ? o : m;
    }
    __shared__ double part[kWaves];
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 0; w < kWaves; ++w) m = part[w] > m ? part[w] : m;
        out[0] = m;
    }
}

// Sum of per-block triples (observing sweeps), fixed order.
__global__ void __launch_bounds__(kBlock)
reduce_final3(const double* partials, int nparts, double* out) {
    double acc[3] = {0, 0, 0};
    for (int p = threadIdx.x; p < nparts; p += kBlock)
        for (int q = 0; q < 3; ++q) acc[q] += partials[3 * p + q];
    block_reduce_store<3>(acc, out);
}

// psi *= s[0], the factor left in device memory by an earlier kernel
// (imaginary time: the deferred 1 / sqrt(norm) of the last step).
template <typename T>
__global__ void __launch_bounds__(kBlock)
scale_state_dev(void* psi_v, long long total, const double* s) {
    cx<T>* psi = static_cast<cx<T>*>(psi_v);
    const double f = *s;
    const long long stride = static_cast<long long>(gridDim.x) * kBlock;
    for (long long f_ = blockIdx.x * static_cast<long long>(kBlock) + threadIdx.x; f_ < total;
         f_ += stride) {
        const cx<T> z = psi[f_];
        psi[f_] = {static_cast<T>(f * z.x), static_cast<T>(f * z.y)};
    }
}

// <a|b> = sum conj(a) b (core/src/field.ixx:88-99): partial sums of the
// real and imaginary parts per block, FP64.
template <typename T>
__global__ void __launch_bounds__(kBlock)
reduce_inner(const void* a_v, const void* b_v, long long total, double* partials) {
    const cx<T>* a = static_cast<const cx<T>*>(a_v);
    const cx<T>* b = static_cast<const cx<T>*>(b_v);
    double acc[kReduceQ] = {};
    const long long stride = static_cast<long long>(gridDim.x) * kBlock;
    for (long long f = blockIdx.x * static_cast<long long>(kBlock) + threadIdx.x; f < total;
         f += stride) {
        const double ar = a[f].x, ai = a[f].y, br = b[f].x, bi = b[f].y;
        acc[0] += ar * br + ai * bi;
        acc[1] += ar * bi - ai * br;
    }
    block_reduce_store<kReduceQ>(acc, partials + blockIdx.x * kReduceQ);
}

// psi -= c phi with c = (sum[0] + i sum[1]) dV from reduce_inner/reduce_final,
// formed in FP64 and rounded once per component.
template <typename T>
__global__ void __launch_bounds__(kBlock)
subtract_projection(void* psi_v, const void* phi_v, long long total, const double* sum, double dv) {
    cx<T>* psi = static_cast<cx<T>*>(psi_v);
    const cx<T>* phi = static_cast<const cx<T>*>(phi_v);
    const double cr = sum[0] * dv, ci = sum[1] * dv;
    const long long stride = static_cast<long long>(gridDim.x) * kBlock;
    for (long long f = blockIdx.x * static_cast<long long>(kBlock) + threadIdx.x; f < total;
         f += stride) {
        const double pr = phi[f].x, pi = phi[f].y;
        const double xr = static_cast<double>(psi[f].x) - (cr * pr - ci * pi);
        const double xi = static_cast<double>(psi[f].y) - (cr * pi + ci * pr);
        psi[f] = {static_cast<T>(xr), static_cast<T>(xi)};
    }
}

// psi /= sqrt(sum[0] dV) (left as is for a zero field, core/src/field.ixx:
// 78-87); the norm is also stored at *norm_out.
template <typename T>
__global__ void __launch_bounds__(kBlock)
normalize_dev(void* psi_v, long long total, const double* sum, double dv, double* norm_out) {
    cx<T>* psi = static_cast<cx<T>*>(psi_v);
    const double n2 = sum[0] * dv;
    const double inv = n2 > 0.0 ? 1.0 / sqrt(n2) : 1.0;
    if (blockIdx.x == 0 && threadIdx.x == 0) *norm_out = n2;
    const long long stride = static_cast<long long>(gridDim.x) * kBlock;
    for (long long f = blockIdx.x * static_cast<long long>(kBlock) + threadIdx.x; f < total;
         f += stride) {
        const cx<T> z = psi[f];
        psi[f] = {static_cast<T>(inv * z.x), static_cast<T>(inv * z.y)};
    }
}

// Energy variance ||(H - <H>) psi||^2 / ||psi||^2 (core/src/observables.ixx:
// 92-119) from psi, T psi (kinetic part, already transformed back) and V:
// pass 1 (mean == null) sums |psi|^2 and Re(conj(psi) H psi) into slots 0
// and 1; pass 2 sums |H psi - <H> psi|^2 into slot 0, with <H> the ratio of
// pass 1's final sums.  H psi = T psi + V psi is formed in FP64.
template <typename T>
__global__ void __launch_bounds__(kBlock)
reduce_hvar(const void* psi_v, const void* tpsi_v, const void* v_v, long long total,
            const double* mean_sums, double* partials) {
    const cx<T>* psi = static_cast<const cx<T>*>(psi_v);
    const cx<T>* tpsi = static_cast<const cx<T>*>(tpsi_v);
    const T* v = static_cast<const T*>(v_v);
    double mean = 0.0;
    if (mean_sums) mean = mean_sums[0] > 0.0 ? mean_sums[1] / mean_sums[0] : 0.0;
    double acc[kReduceQ] = {};
    const long long stride = static_cast<long long>(gridDim.x) * kBlock;
    for (long long f = blockIdx.x * static_cast<long long>(kBlock) + threadIdx.x; f < total;
         f += stride) {
        const double pr = psi[f].x, pi = psi[f].y;
        const double vf = v ? static_cast<double>(v[f]) : 0.0;
        const double hr = static_cast<double>(tpsi[f].x) + vf * pr;
        const double hi = static_cast<double>(tpsi[f].y) + vf * pi;
        if (!mean_sums) {
            acc[0] += pr * pr + pi * pi;
            acc[1] += pr * hr + pi * hi;
        } else {
            const double dr = hr - mean * pr, di = hi - mean * pi;
            acc[0] += dr * dr + di * di;
        }
    }
    block_reduce_store<kReduceQ>(acc, partials + blockIdx.x * kReduceQ);
}

// Unfused FP64 elementwise passes of the reference data flow (the baseline
// the fused path is compared with): psi *= table, psi = conj(psi) * s,
// psi *= mask.
__global__ void __launch_bounds__(kBlock)
table_mul(void* psi_v, const void* tab_v, long long total) {
    cx<double>* psi = static_cast<cx<double>*>(psi_v);
    const cx<double>* tab = static_cast<const cx<double>*>(tab_v);
    const long long stride = static_cast<long long>(gridDim.x) * kBlock;
    for (long long f = blockIdx.x * static_cast<long long>(kBlock) + threadIdx.x; f < total;
         f += stride)
        psi[f] = cmul(psi[f], tab[f]);
}

__global__ void __launch_bounds__(kBlock)
conj_scale(void* psi_v, long long total, double s) {
    cx<double>* psi = static_cast<cx<double>*>(psi_v);
    const long long stride = static_cast<long long>(gridDim.x) * kBlock;
    for (long long f = blockIdx.x * static_cast<long long>(kBlock) + threadIdx.x; f < total;
         f += stride) {
        const cx<double> z = psi[f];
        psi[f] = {s * z.x, s * -z.y};
    }
}

__global__ void __launch_bounds__(kBlock)
mask_mul(void* psi_v, const double* mask, long long total) {
    cx<double>* psi = static_cast<cx<double>*>(psi_v);
    const long long stride = static_cast<long long>(gridDim.x) * kBlock;
    for (long long f = blockIdx.x * static_cast<long long>(kBlock) + threadIdx.x; f < total;
         f += stride)
        psi[f] = cscale(psi[f], mask[f]);
}

// A plain streaming copy, 16 bytes per lane per load and store, grid-stride:
// the bandwidth reference the sweeps are compared with (a kernel like theirs,
// reading and writing each byte once).
__global__ void __launch_bounds__(kBlock) copy16(uint4* __restrict__ dst,
                                                 const uint4* __restrict__ src, long long n16) {
    const long long stride = static_cast<long long>(gridDim.x) * kBlock;
    long long i = blockIdx.x * static_cast<long long>(kBlock) + threadIdx.x;
    for (; i + 3 * stride < n16; i += 4 * stride) {
        const uint4 a = src[i], b = src[i + stride], c = src[i + 2 * stride], d = src[i + 3 * stride];
        dst[i] = a;
        dst[i + stride] = b;
        dst[i + 2 * stride] = c;
        dst[i + 3 * stride] = d;
    }
    for (; i < n16; i += stride) dst[i] = src[i];
}

int elementwise_blocks(long long total) {
    long long b = (total + kBlock * 4 - 1) / (kBlock * 4);
    if (b < 1) b = 1;
    if (b > 16384) b = 16384;
    return static_cast<int>(b);
}

}  // namespace

hipError_t launch_table_mul(void* psi, const void* table, long long total, hipStream_t stream) {
    hipLaunchKernelGGL(table_mul, dim3(elementwise_blocks(total)), dim3(kBlock), 0, stream, psi,
                       table, total);
    return hipGetLastError();
}

hipError_t launch_conj_scale(void* psi, long long total, double s, hipStream_t stream) {
    hipLaunchKernelGGL(conj_scale, dim3(elementwise_blocks(total)), dim3(kBlock), 0, stream, psi,
                       total, s);
    return hipGetLastError();
}

hipError_t launch_mask_mul(void* psi, const double* mask, long long total, hipStream_t stream) {
    hipLaunchKernelGGL(mask_mul, dim3(elementwise_blocks(total)), dim3(kBlock), 0, stream, psi,
                       mask, total);
    return hipGetLastError();
}

hipError_t launch_copy(void* dst, const void* src, size_t bytes, hipStream_t stream) {
    const long long n16 = static_cast<long long>(bytes / 16);
    if (n16 > 0) {
        long long blocks = (n16 + kBlock * 8 - 1) / (kBlock * 8);
        if (blocks > 8192) blocks = 8192;
        hipLaunchKernelGGL(copy16, dim3(static_cast<unsigned>(blocks)), dim3(kBlock), 0, stream,
                           static_cast<uint4*>(dst), static_cast<const uint4*>(src), n16);
    }
    const size_t tail = bytes - static_cast<size_t>(n16) * 16;
    if (tail)
        return hipMemcpyAsync(static_cast<char*>(dst) + n16 * 16,
                              static_cast<const char*>(src) + n16 * 16, tail,
                              hipMemcpyDeviceToDevice, stream);
    return hipGetLastError();
}

hipError_t launch_peak(bool fp64, const void* psi, long long total, double* partials, double* out,
                       hipStream_t stream) {
    const int blocks = reduce_blocks(total);
    if (fp64)
        hipLaunchKernelGGL(reduce_peak<double>, dim3(blocks), dim3(kBlock), 0, stream, psi, total,
                           partials);
    else
        hipLaunchKernelGGL(reduce_peak<float>, dim3(blocks), dim3(kBlock), 0, stream, psi, total,
                           partials);
    hipLaunchKernelGGL(reduce_peak_final, dim3(1), dim3(kBlock), 0, stream, partials, blocks, out);
    return hipGetLastError();
}

int reduce_blocks(long long total) {
    long long b = (total + kBlock * 8 - 1) / (kBlock * 8);
    if (b < 1) b = 1;
    if (b > kMaxReduceBlocks) b = kMaxReduceBlocks;
    return static_cast<int>(b);
}

hipError_t launch_reduce(bool fp64, bool kspace, const ReduceArgs& a, double* partials,
                         double* out, hipStream_t stream) {
    const int blocks = reduce_blocks(a.total);
    if (kspace) {
        if (fp64)
            hipLaunchKernelGGL(reduce_k<double>, dim3(blocks), dim3(kBlock), 0, stream, a, partials);
        else
            hipLaunchKernelGGL(reduce_k<float>, dim3(blocks), dim3(kBlock), 0, stream, a, partials);
    } else {
        if (fp64)
            hipLaunchKernelGGL(reduce_x<double>, dim3(blocks), dim3(kBlock), 0, stream, a, partials);
        else
            hipLaunchKernelGGL(reduce_x<float>, dim3(blocks), dim3(kBlock), 0, stream, a, partials);
    }
    hipLaunchKernelGGL(reduce_final, dim3(1), dim3(kBlock), 0, stream, partials, blocks, out);
    return hipGetLastError();
}

namespace {
long long scale_blocks(long long total) {
    long long b = (total + kBlock * 4 - 1) / (kBlock * 4);
    if (b < 1) b = 1;
    if (b > 8192) b = 8192;
    return b;
}
}  // namespace

hipError_t launch_energy_variance(bool fp64, const void* psi, const void* tpsi, const void* v,
                                 long long total, double* partials, double* out,
                                 hipStream_t stream) {
    const int blocks = reduce_blocks(total);
    double* first = out;             // pass 1 sums (|psi|^2, Re <psi|H psi>)
    double* second = out + kReduceQ;  // pass 2 sum |H psi - <H> psi|^2
    for (int pass = 0; pass < 2; ++pass) {
        const double* mean = pass ? first : nullptr;
        if (fp64)
            hipLaunchKernelGGL(reduce_hvar<double>, dim3(blocks), dim3(kBlock), 0, stream, psi, tpsi,
                               v, total, mean, partials);
        else
            hipLaunchKernelGGL(reduce_hvar<float>, dim3(blocks), dim3(kBlock), 0, stream, psi, tpsi,
                               v, total, mean, partials);
        hipLaunchKernelGGL(reduce_final, dim3(1), dim3(kBlock), 0, stream, partials, blocks,
                           pass ? second : first);
    }
    return hipGetLastError();
}

hipError_t launch_reduce_final3(const double* partials, int nblocks, double* out,
                                hipStream_t stream) {
    hipLaunchKernelGGL(reduce_final3, dim3(1), dim3(kBlock), 0, stream, partials, nblocks, out);
    return hipGetLastError();
}

hipError_t launch_scale_dev(bool fp64, void* psi, long long total, const double* s,
                            hipStream_t stream) {
    const long long b = scale_blocks(total);
    if (fp64)
        hipLaunchKernelGGL(scale_state_dev<double>, dim3(b), dim3(kBlock), 0, stream, psi, total, s);
    else
        hipLaunchKernelGGL(scale_state_dev<float>, dim3(b), dim3(kBlock), 0, stream, psi, total, s);
    return hipGetLastError();
}

hipError_t launch_inner(bool fp64, const void* a, const void* b, long long total, double* partials,
                        double* out, hipStream_t stream) {
    const int blocks = reduce_blocks(total);
    if (fp64)
        hipLaunchKernelGGL(reduce_inner<double>, dim3(blocks), dim3(kBlock), 0, stream, a, b, total,
                           partials);
    else
        hipLaunchKernelGGL(reduce_inner<float>, dim3(blocks), dim3(kBlock), 0, stream, a, b, total,
                           partials);
    hipLaunchKernelGGL(reduce_final, dim3(1), dim3(kBlock), 0, stream, partials, blocks, out);
    return hipGetLastError();
}

hipError_t launch_subtract_projection(bool fp64, void* psi, const void* phi, long long total,
                                      const double* sum, double dv, hipStream_t stream) {
    const long long b = scale_blocks(total);
    if (fp64)
        hipLaunchKernelGGL(subtract_projection<double>, dim3(b), dim3(kBlock), 0, stream, psi, phi,
                           total, sum, dv);
    else
        hipLaunchKernelGGL(subtract_projection<float>, dim3(b), dim3(kBlock), 0, stream, psi, phi,
                           total, sum, dv);
    return hipGetLastError();
}

hipError_t launch_normalize_dev(bool fp64, void* psi, long long total, const double* sum, double dv,
                                double* norm_out, hipStream_t stream) {
    const long long b = scale_blocks(total);
    if (fp64)
        hipLaunchKernelGGL(normalize_dev<double>, dim3(b), dim3(kBlock), 0, stream, psi, total, sum,
                           dv, norm_out);
    else
        hipLaunchKernelGGL(normalize_dev<float>, dim3(b), dim3(kBlock), 0, stream, psi, total, sum,
                           dv, norm_out);
    return hipGetLastError();
}

hipError_t launch_scale(bool fp64, void* psi, long long total, double s, hipStream_t stream) {
    long long b = (total + kBlock * 4 - 1) / (kBlock * 4);
    if (b < 1) b = 1;
    if (b > 8192) b = 8192;
    if (fp64)
        hipLaunchKernelGGL(scale_state<double>, dim3(b), dim3(kBlock), 0, stream, psi, total, s);
    else
        hipLaunchKernelGGL(scale_state<float>, dim3(b), dim3(kBlock), 0, stream, psi, total, s);
    return hipGetLastError();
}

// Imaginary time: the norm of a measuring sweep from its per-block partials
// (block order, fixed-order tree), then scale[parity ^ 1] = 1 / sqrt(norm)
// (a zero field is left as it is, core/src/field.ixx:78-81), norm and peak.
__global__ void __launch_bounds__(256) relax_finish(const double* partials, RelaxSlots* slots,
                                                    int parity, double cell) {
    relax_finish_block(partials, slots, parity, cell);
}

hipError_t launch_relax_finish(const SweepArgs& a, int, hipStream_t stream) {
    hipLaunchKernelGGL(relax_finish, dim3(1), dim3(256), 0, stream, a.partials, a.relax, a.parity,
                       a.cell);
    return hipGetLastError();
}

}  // namespace tdse
