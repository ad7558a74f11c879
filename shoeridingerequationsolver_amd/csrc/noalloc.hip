// noalloc.hip -- tdse_observe_noalloc: the position-space pass of
// tdse_observe over a state the caller holds in device memory, for callers
// that cannot allocate device memory (code under HIP graph capture, serving
// loops, harnesses that forbid hipMalloc).
//
// The reduction's per-block partials and its final sums live in `scratch`, a
// file-scope device array: each device holds its own copy from the time the
// code object is loaded until the process ends.  Its device address is taken
// once per device with hipGetSymbolAddress (the array's name in host code is
// only a host-side stand-in) and that address is what the kernels receive.
// Every call on one device shares that device's array: calls that overlap on
// one device (different streams, different host threads without ordering)
// race on it.  Each block writes its own partial and the final pass writes
// every sum before the finishing kernel reads them, so nothing needs clearing
// between calls.
#include <hip/hip_runtime.h>

#include <cmath>
#include <mutex>

#include "../../include/tdse.h"
#include "launch.hpp"
#include "observe.hpp"

namespace {

// partials of kMaxReduceBlocks blocks (kReduceQ sums each), then the kReduceQ
// final sums: 2048 * 9 + 9 doubles = 147,528 bytes per device
constexpr int kScratchDoubles = tdse::kMaxReduceBlocks * tdse::kReduceQ + tdse::kReduceQ;
static __device__ __align__(256) double scratch[kScratchDoubles];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
double* g_ptr[kMaxDevices] = {};  // device address of `scratch`, per device

// The ratios of tdse_observe (same expressions, FP64, IEEE division and
// square root) from the final sums: |psi|^2, x|psi|^2, y|psi|^2, z|psi|^2,
// V|psi|^2, slab, x^2|psi|^2, y^2|psi|^2, z^2|psi|^2.
__device__ __forceinline__ double ratio(double num, double den) { return den > 0.0 ? num / den : 0.0; }

__global__ void __launch_bounds__(64) finish_observables(const double* sums, double cell,
                                                         tdse_observables* out) {
    if (threadIdx.x != 0) return;
    const double s0 = sums[0];
    tdse_observables o;
    o.sum_abs2 = s0;
    o.norm = s0 * cell;
    for (int a = 0; a < 3; ++a) o.position[a] = ratio(sums[1 + a], s0);
    o.potential = ratio(sums[4], s0);
    for (int a = 0; a < 3; ++a) {
        const double m = ratio(sums[1 + a], s0);
        o.sigma[a] = sqrt(fmax(0.0, ratio(sums[6 + a], s0) - m * m));
        o.momentum[a] = 0.0;
    }
    o.kinetic = 0.0;
    o.energy = 0.0;
    o.energy_variance = 0.0;
    o.peak = 0.0;
    *out = o;
}

int fail(int status, const char* what) { return tdse::record_error(status, what); }

// The device address of this device's `scratch` (taken once per device, with
// that device current).
int scratch_ptr(double** out) {
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) {
        (void)hipGetLastError();
        return fail(TDSE_E_NO_DEVICE, "no current HIP device");
    }
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_ptr[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            (void)hipGetLastError();
            return fail(TDSE_E_DEVICE, "hipGetSymbolAddress of the scratch array failed");
        }
        g_ptr[dev] = static_cast<double*>(p);
    }
    *out = g_ptr[dev];
    return TDSE_OK;
}

}  // namespace

extern "C" {

int tdse_observe_noalloc(const void* psi, const void* v, int precision, const int n[3],
                         const double lo[3], const double hi[3], tdse_observables* out,
                         void* stream) {
    if (!psi || !n || !lo || !hi || !out) return fail(TDSE_E_INVALID, "null argument");
    if (precision != TDSE_FP32 && precision != TDSE_FP64)
        return fail(TDSE_E_INVALID, "precision must be TDSE_FP32 or TDSE_FP64");
    long long total = 1;
    for (int a = 0; a < 3; ++a) {
        if (n[a] < 1) return fail(TDSE_E_INVALID, "grid extent < 1");
        if (!(hi[a] > lo[a]) && n[a] > 1) return fail(TDSE_E_INVALID, "hi must exceed lo");
        total *= n[a];
    }
    // the largest state a plan holds (32-bit buffer offsets in the sweeps)
    if (total * (precision == TDSE_FP64 ? 16 : 8) >= (1LL << 32))
        return fail(TDSE_E_UNSUPPORTED, "state larger than 4 GiB");
    double* sp = nullptr;
    int rc = scratch_ptr(&sp);
    if (rc) return rc;
    tdse::ReduceArgs r{};
    r.psi = psi;
    r.v = v;
    r.total = total;
    for (int a = 0; a < 3; ++a) {
        r.n[a] = n[a];
        r.lo[a] = lo[a];
        r.h[a] = (hi[a] - lo[a]) / n[a];
        r.sq[a] = 1;
        r.dk[a] = 0.0;
    }
    tdse::empty_box(r);
    r.mass = 1.0;
    // same block count, partial layout and final tree as tdse_observe
    static_assert(kScratchDoubles >= tdse::kMaxReduceBlocks * tdse::kReduceQ + tdse::kReduceQ,
                  "scratch holds the partials of the most blocks a reduction launches");
    double* partials = sp;
    double* sums = sp + tdse::kMaxReduceBlocks * tdse::kReduceQ;
    hipStream_t st = static_cast<hipStream_t>(stream);
    if (tdse::launch_reduce(precision == TDSE_FP64, false, r, partials, sums, st) != hipSuccess)
        return fail(TDSE_E_DEVICE, "reduction launch failed");
    const double cell = r.h[0] * r.h[1] * r.h[2];
    hipLaunchKernelGGL(finish_observables, dim3(1), dim3(64), 0, st, sums, cell, out);
    if (hipGetLastError() != hipSuccess) return fail(TDSE_E_DEVICE, "finish launch failed");
    return TDSE_OK;
}

}  // extern "C"
