// tdse_api.cpp -- the C ABI of include/tdse.h: contexts, plans, tables,
// the sweep schedule of a Strang step, observables and timing.
//
// Replaces ses::SplitOperator1D/3D (core/src/propagator.ixx:21-134) for a
// caller that owns psi (core/src/field.ixx layout) and wants it stepped on the
// GPU.  Everything that runs per step is a HIP kernel; the host only builds
// the O(N) tables once (as the reference constructor does) and enqueues.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/tdse.h"
#include "launch.hpp"
#include "observe.hpp"
#include "device_build.hpp"
#include "lattice.hpp"

using tdse::SweepArgs;

namespace {

thread_local std::string g_last_error;

int fail(int status, const std::string& what) {
    g_last_error = what;
    return status;
}

#define TDSE_HIP(call)                                                                     \
    do {                                                                                   \
        hipError_t e_ = (call);                                                            \
        if (e_ != hipSuccess)                                                              \
            return fail(TDSE_E_DEVICE, std::string(#call ": ") + hipGetErrorString(e_));   \
    } while (0)

constexpr double kTwoPiHi = 6.28318530717958623200e+00;
constexpr double kTwoPiLo = 2.44929359829470635445e-16;

// FP64 reduction of an angle into [-pi, pi] (two-term 2 pi, exact n * hi).
double reduce_2pi_host(double th) {
    const double n = std::nearbyint(th / (2.0 * 3.14159265358979323846));
    return (th - n * kTwoPiHi) - n * kTwoPiLo;
}

struct Axis {
    int n;
    double lo, hi;
    double length() const { return hi - lo; }
};

enum Kind { kKind3D, kKind2D, kKindSingle, kKindFourStep, kKindLattice };

constexpr int kMaxSlots = 6;

// Launch::mode of the small kernel that ends a measuring sweep
constexpr int kFinishMode = -1;

struct Launch {
    int mode;
    int n;          // line length
    SweepArgs args;
    tdse::LaunchFn fn;
    int slot;       // profiling slot
};

}  // namespace

int tdse::record_error(int status, const char* what) { return fail(status, what); }

thread_local tdse::LaunchTiming tdse::g_launch_timing;

// Two pinned host buffers per context for state uploads and downloads: the
// host conversion of one chunk overlaps the DMA of the other.  Allocated on
// the first transfer; without them transfers go through pageable memory.
struct PinnedStage {
    static constexpr size_t kBytes = size_t(32) << 20;
    void* buf[2] = {nullptr, nullptr};
    hipEvent_t ev[2] = {nullptr, nullptr};
    bool tried = false;
    std::mutex mu;  // one transfer at a time per context
};

struct tdse_ctx {
    int device = 0;
    uint64_t budget = 0;
    uint64_t used = 0;
    int max_blocks = 1 << 20;
    std::mutex mu;
    PinnedStage pin;
    // streams of the context's live plans and batches (tdse_sync waits on
    // them); reg_mu is held for a whole tdse_sync, so a plan or batch is not
    // destroyed while its stream is waited on
    std::mutex reg_mu;
    std::vector<hipStream_t> streams;
};

struct tdse_plan {
    tdse_ctx* ctx = nullptr;
    int precision = TDSE_FP32;
    int n[3] = {1, 1, 1};
    double lo[3] = {0, 0, 0}, hi[3] = {1, 1, 1};
    double dt = 0, mass = 1;
    long long total = 0;
    std::vector<Axis> sq;  // axes with n > 1, x-fastest order
    Kind kind = kKind3D;
    int n1 = 0, n2 = 0;    // four-step split: rows of n1 contiguous points, n2 rows
    hipStream_t stream = nullptr;
    void* psi = nullptr;
    void* phi = nullptr;
    void* v = nullptr;
    void* mask = nullptr;
    // a mask that is a product of per-axis ramps (the reference's absorbing
    // mask): the x ramp in stored form and the y, z ramps in FP64, read by
    // the x sweeps instead of the full array (which the other sweeps keep)
    void* mask_x = nullptr;
    double* mask_yz = nullptr;
    void* scratch = nullptr;  // forward transform for <T>; T psi for Var(H)
    void* kvar = nullptr;     // 1D: (0.5 k_j^2 / m, 0) per bin, FP64 (Var(H))
    double* partials = nullptr;
    double* red_out = nullptr;
    std::map<int, void*> tw;  // line length -> W_N table
    void* kt[3] = {nullptr, nullptr, nullptr};  // per squeezed axis exp(-i k^2 dt / 2m), FP64
    long long kt_len[3] = {0, 0, 0};
    bool has_potential = false;
    uint64_t bytes = 0;
    unsigned flags = 0;
    int variant[kMaxSlots] = {0, 0, 0, 0, 0, 0};  // kernel variant per schedule slot
    // dipole drive V = amplitude (axis . r) cos(omega t) (core/src/drive.ixx:19-23)
    bool drive = false;
    double drive_axis[3] = {0, 0, 0};
    double drive_amp = 0.0, drive_omega = 0.0;
    double time = 0.0;  // plan clock: advanced by dt per step
    // uniform magnetic field along axis b_axis (core/src/magnetic.ixx:19-80)
    double bfield = 0.0;
    int b_axis = 2;
    // Peierls lattice x links (core/src/lattice2d.ixx:63-105): 0 none,
    // 1 uniform field (angle lat_a per unit y), 2 solenoid (flux lat_a on
    // column lat_col, rows [lat_j0, lat_j1), sign lat_sign)
    int lat_mode = 0;
    double lat_a = 0.0, lat_sign = -1.0;
    int lat_col = -1, lat_j0 = 0, lat_j1 = 0;
    double hop(int axis) const {
        const double h = (hi[axis] - lo[axis]) / n[axis];
        return 0.5 / (h * h);
    }
    // unfused FP64 baseline: full complex128 tables of the reference propagator
    void* base_half = nullptr;  // exp(-i V dt / 2) per cell
    void* base_kin = nullptr;   // exp(-i |k|^2 dt / 2m) per cell

    size_t csize() const { return precision == TDSE_FP64 ? 16 : 8; }
    size_t rsize() const { return precision == TDSE_FP64 ? 8 : 4; }
    bool fp64() const { return precision == TDSE_FP64; }
    bool imag() const { return (flags & TDSE_FLAG_IMAGINARY_TIME) != 0; }
    // imaginary time: device slots inside the reduction buffer (after red_out)
    tdse::RelaxSlots* relax() const { return reinterpret_cast<tdse::RelaxSlots*>(red_out + 16); }
    // Real-time sweep plans store the kick phase phi as tan(phi / 2): the
    // kernels form exp(i phi) = (1 - t^2 + 2 i t) / (1 + t^2) without a sine
    // or cosine (the lattice keeps phi)
    bool tan_kick() const { return !imag() && kind != kKindLattice; }
};

namespace {

tdse::LaunchFn launcher(const tdse_plan* p, int n, int mode, int variant = 0) {
    if (variant != 0)
        return p->fp64() ? tdse::find_variant_launcher_f64(n, mode, variant)
                         : tdse::find_variant_launcher_f32(n, mode, variant);
    return p->fp64() ? tdse::find_launcher_f64(n, mode) : tdse::find_launcher_f32(n, mode);
}

int set_device(const tdse_ctx* ctx) {
    TDSE_HIP(hipSetDevice(ctx->device));
    return TDSE_OK;
}

void register_stream(tdse_ctx* ctx, hipStream_t s) {
    std::lock_guard<std::mutex> lk(ctx->reg_mu);
    ctx->streams.push_back(s);
}

void unregister_stream(tdse_ctx* ctx, hipStream_t s) {
    std::lock_guard<std::mutex> lk(ctx->reg_mu);
    ctx->streams.erase(std::remove(ctx->streams.begin(), ctx->streams.end(), s),
                       ctx->streams.end());
}

// Four-step split N = n1 * n2 (n1 contiguous row length, n2 rows), both
// compiled lengths; prefer the most balanced split, then the longer rows.
bool four_step_split(long long n, int* n1, int* n2) {
    int best1 = 0, best2 = 0;
    for (int a = 2; a <= 1024; ++a) {
        if (!tdse::line_supported(a) || n % a) continue;
        const long long b = n / a;
        if (b > 1024 || !tdse::line_supported(static_cast<int>(b))) continue;
        const int mn = std::min<long long>(a, b), bmn = std::min(best1, best2);
        if (mn > bmn || (mn == bmn && a > best1)) {
            best1 = a;
            best2 = static_cast<int>(b);
        }
    }
    if (!best1) return false;
    *n1 = best1;
    *n2 = best2;
    return true;
}

// Device bytes of a plan: state, kick phase, potential, mask array, line
// twiddles, kinetic tables, reduction partials and, for a masked grid whose x
// sweeps can read per-axis ramps (3D, or 2D with x lines along the first
// axis), the ramps (x in the plan's precision, y and z in FP64).
uint64_t plan_bytes_for(const int n[3], long long total, const std::vector<int>& lines, bool fp64,
                        bool mask, long long kt_points, bool ramps) {
    const uint64_t cs = fp64 ? 16 : 8, rs = fp64 ? 8 : 4;
    uint64_t b = total * cs + 2 * total * rs + (mask ? total * rs : 0);
    for (int len : lines) b += static_cast<uint64_t>(len) * 16;
    b += static_cast<uint64_t>(kt_points) * 16;
    b += static_cast<uint64_t>(tdse::kMaxReduceBlocks) * tdse::kReduceQ * 8 + 64 * 8;
    if (mask && ramps) b += static_cast<uint64_t>(n[0]) * rs + static_cast<uint64_t>(n[1] + n[2]) * 8;
    return b;
}

int classify(const int n[3], std::vector<Axis>& sq, const double lo[3], const double hi[3],
             Kind* kind, int* n1, int* n2, std::vector<int>* lines) {
    sq.clear();
    for (int a = 0; a < 3; ++a)
        if (n[a] > 1) sq.push_back({n[a], lo ? lo[a] : 0.0, hi ? hi[a] : 1.0});
    lines->clear();
    if (sq.empty()) return fail(TDSE_E_UNSUPPORTED, "grid of one point has no FFT axis");
    if (sq.size() == 1) {
        const int N = sq[0].n;
        if (N <= 1024 && tdse::line_supported(N)) {
            *kind = kKindSingle;
            lines->push_back(N);
            return TDSE_OK;
        }
        if (!four_step_split(N, n1, n2))
            return fail(TDSE_E_UNSUPPORTED, "1D length " + std::to_string(N) +
                                                " is not a product of two compiled line lengths");
        *kind = kKindFourStep;
        lines->push_back(*n1);
        if (*n2 != *n1) lines->push_back(*n2);
        return TDSE_OK;
    }
    for (const Axis& ax : sq) {
        if (!tdse::line_supported(ax.n))
            return fail(TDSE_E_UNSUPPORTED, "line length " + std::to_string(ax.n) +
                                                " has no compiled FFT plan");
        if (std::find(lines->begin(), lines->end(), ax.n) == lines->end()) lines->push_back(ax.n);
    }
    *kind = sq.size() == 2 ? kKind2D : kKind3D;
    return TDSE_OK;
}

// Host <-> device copies ordered on the plan's stream, complete on return
// (the host buffer may then be reused).  Plan streams are non-blocking, so
// nothing queued on them waits for null-stream copies or memsets: every
// host-initiated write of plan memory goes through the plan's stream.
int h2d(const tdse_plan* p, void* dst, const void* src, size_t bytes) {
    TDSE_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, p->stream));
    TDSE_HIP(hipStreamSynchronize(p->stream));
    return TDSE_OK;
}

int d2h(const tdse_plan* p, void* dst, const void* src, size_t bytes) {
    TDSE_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, p->stream));
    TDSE_HIP(hipStreamSynchronize(p->stream));
    return TDSE_OK;
}

// f(i0, i1) over [0, n) in contiguous chunks on up to 16 host threads (one
// per 2^18 elements): the element-wise conversions between the
// caller's FP64 arrays and the plan's storage, each element independent, so
// the result does not depend on the split.
template <class F>
void parallel_chunks(long long n, F&& f) {
    const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
    const long long want = std::min<long long>(std::min<unsigned>(hw, 16u), n >> 18);
    if (want <= 1) {
        f(0LL, n);
        return;
    }
    std::vector<std::thread> pool;
    const long long per = (n + want - 1) / want;
    for (long long t = 1; t < want; ++t) {
        const long long a = t * per, b = std::min(n, a + per);
        if (a >= b) continue;
        try {
            pool.emplace_back([&f, a, b] { f(a, b); });
        } catch (...) {  // no thread to be had: this one does the chunk
            f(a, b);
        }
    }
    f(0LL, std::min(n, per));
    for (auto& th : pool) th.join();
}

enum RealForm { kPlain, kPhase, kTanHalf, kMask };

// A per-cell real array in the plan's precision: plain values, kick phases
// reduced to [-pi, pi], the tangent of half such a phase, or an absorbing
// mask (stored as 1 - m in FP32 plans, m in FP64 plans; the kernels undo it
// in FP64).
template <typename T>
std::vector<T> to_real(const double* src, long long n, RealForm form) {
    std::vector<T> out(static_cast<size_t>(n));
    parallel_chunks(n, [&](long long i0, long long i1) {
        for (long long i = i0; i < i1; ++i) {
            double x = src[i];
            if (form == kPhase) x = reduce_2pi_host(x);
            else if (form == kTanHalf) x = std::tan(0.5 * reduce_2pi_host(x));
            else if (form == kMask && sizeof(T) == 4) x = 1.0 - x;
            out[i] = static_cast<T>(x);
        }
    });
    return out;
}

// The context's pinned buffers, allocated on first use (false: none to be had).
bool pinned_stage(tdse_ctx* ctx) {
    PinnedStage& st = ctx->pin;
    if (!st.tried) {
        st.tried = true;
        bool ok = true;
        for (int b = 0; b < 2 && ok; ++b)
            ok = hipHostMalloc(&st.buf[b], PinnedStage::kBytes, hipHostMallocDefault) == hipSuccess &&
                 hipEventCreateWithFlags(&st.ev[b], hipEventDisableTiming) == hipSuccess;
        if (!ok) {
            (void)hipGetLastError();
            for (int b = 0; b < 2; ++b) {
                if (st.buf[b]) (void)hipHostFree(st.buf[b]);
                if (st.ev[b]) (void)hipEventDestroy(st.ev[b]);
                st.buf[b] = nullptr;
                st.ev[b] = nullptr;
            }
        }
    }
    return st.buf[0] != nullptr;
}

// n FP64 host values -> device array of T, through the pinned buffers in
// chunks: chunk k + 1 is converted while chunk k is in flight.
template <typename T>
int upload_staged(tdse_plan* p, void* dst, const double* src, long long n) {
    PinnedStage& st = p->ctx->pin;
    const long long per = static_cast<long long>(PinnedStage::kBytes / sizeof(T));
    for (long long off = 0, k = 0; off < n; off += per, ++k) {
        const long long len = std::min(per, n - off);
        const int b = static_cast<int>(k & 1);
        TDSE_HIP(hipEventSynchronize(st.ev[b]));  // the buffer's previous DMA is done
        T* h = static_cast<T*>(st.buf[b]);
        parallel_chunks(len, [&](long long i0, long long i1) {
            for (long long i = i0; i < i1; ++i) h[i] = static_cast<T>(src[off + i]);
        });
        TDSE_HIP(hipMemcpyAsync(static_cast<T*>(dst) + off, h, len * sizeof(T),
                                hipMemcpyHostToDevice, p->stream));
        TDSE_HIP(hipEventRecord(st.ev[b], p->stream));
    }
    TDSE_HIP(hipStreamSynchronize(p->stream));
    return TDSE_OK;
}

// Device array of T (n values) -> FP64 host values: the DMA of chunk k + 1 is
// in flight while chunk k is converted.
template <typename T, typename D = double>
int download_staged(tdse_plan* p, D* dst, const void* src, long long n) {
    PinnedStage& st = p->ctx->pin;
    const long long per = static_cast<long long>(PinnedStage::kBytes / sizeof(T));
    const long long chunks = (n + per - 1) / per;
    auto issue = [&](long long k) {
        const long long off = k * per, len = std::min(per, n - off);
        TDSE_HIP(hipMemcpyAsync(st.buf[k & 1], static_cast<const T*>(src) + off, len * sizeof(T),
                                hipMemcpyDeviceToHost, p->stream));
        TDSE_HIP(hipEventRecord(st.ev[k & 1], p->stream));
        return TDSE_OK;
    };
    int rc = chunks > 0 ? issue(0) : TDSE_OK;
    for (long long k = 0; k < chunks && !rc; ++k) {
        // chunk k + 1 goes into the buffer chunk k - 1 was converted from
        if (k + 1 < chunks && (rc = issue(k + 1))) break;
        TDSE_HIP(hipEventSynchronize(st.ev[k & 1]));
        const long long off = k * per, len = std::min(per, n - off);
        const T* h = static_cast<const T*>(st.buf[k & 1]);
        parallel_chunks(len, [&](long long i0, long long i1) {
            for (long long i = i0; i < i1; ++i) dst[off + i] = static_cast<D>(h[i]);
        });
    }
    return rc;
}

int upload_real(tdse_plan* p, void* dst, const double* src, RealForm form) {
    if (p->fp64()) {
        std::vector<double> h = to_real<double>(src, p->total, form);
        return h2d(p, dst, h.data(), h.size() * 8);
    }
    std::vector<float> h = to_real<float>(src, p->total, form);
    return h2d(p, dst, h.data(), h.size() * 4);
}

int alloc(tdse_plan* p, void** out, uint64_t bytes) {
    {
        std::lock_guard<std::mutex> lk(p->ctx->mu);
        if (p->ctx->budget && p->ctx->used + bytes > p->ctx->budget)
            return fail(TDSE_E_OOM, "HBM budget exceeded");
        p->ctx->used += bytes;
    }
    if (hipMalloc(out, bytes) != hipSuccess) {
        (void)hipGetLastError();
        std::lock_guard<std::mutex> lk(p->ctx->mu);
        p->ctx->used -= bytes;
        *out = nullptr;
        return fail(TDSE_E_OOM, "hipMalloc of " + std::to_string(bytes) + " bytes failed");
    }
    p->bytes += bytes;
    return TDSE_OK;
}

void release(tdse_plan* p, void*& ptr, uint64_t bytes) {
    if (!ptr) return;
    (void)hipFree(ptr);
    ptr = nullptr;
    std::lock_guard<std::mutex> lk(p->ctx->mu);
    p->ctx->used -= bytes;
    p->bytes -= bytes;
}

// Twiddle table W_N^m = exp(-2 pi i m / N), m < N, FP64 for both state
// precisions (the kernels form and apply twiddles in FP64).
int make_twiddles(tdse_plan* p, int N) {
    if (p->tw.count(N)) return TDSE_OK;
    void* d = nullptr;
    int rc = alloc(p, &d, static_cast<uint64_t>(N) * 16);
    if (rc) return rc;
    p->tw[N] = d;
    std::vector<double> w(2 * static_cast<size_t>(N));
    for (int m = 0; m < N; ++m) {
        const double th = -2.0 * 3.14159265358979323846 * static_cast<double>(m) / N;
        w[2 * m] = std::cos(th);
        w[2 * m + 1] = std::sin(th);
    }
    return h2d(p, d, w.data(), w.size() * 8);
}

// Kinetic phase table of one axis (or, for a four-step 1D grid, of the whole
// line): exp(-i * 0.5 * k_j^2 / mass * dt) with k_j the fftfreq wavenumber
// (core/src/spectral.ixx:14-23, core/src/propagator.ixx:34-37, 83-98), FP64.
int make_kinetic(tdse_plan* p, int slot, const Axis& ax) {
    void* d = nullptr;
    int rc = alloc(p, &d, static_cast<uint64_t>(ax.n) * 16);
    if (rc) return rc;
    p->kt[slot] = d;
    p->kt_len[slot] = ax.n;
    const double dk = 2.0 * 3.14159265358979323846 / (ax.hi - ax.lo);
    std::vector<double> k(2 * static_cast<size_t>(ax.n));
    for (int j = 0; j < ax.n; ++j) {
        const double kj = dk * ((2 * j < ax.n) ? j : j - ax.n);
        const double th = -0.5 * kj * kj / p->mass * p->dt;
        if (p->imag()) {
            // imaginary time: the real weight exp(-k^2 dtau / 2m)
            // (core/src/imaginary_time.ixx:35-39, 82-93)
            k[2 * j] = std::exp(th);
            k[2 * j + 1] = 0.0;
        } else {
            k[2 * j] = std::cos(th);
            k[2 * j + 1] = std::sin(th);
        }
    }
    return h2d(p, d, k.data(), k.size() * 8);
}

// The reference SplitOperator3D tables (core/src/propagator.ixx:72-101):
// full N-entry complex128 kinetic table over the three original axes.
int make_baseline_kinetic(tdse_plan* p) {
    std::vector<double> k[3];
    for (int a = 0; a < 3; ++a) {
        k[a].resize(p->n[a]);
        const double dk = 2.0 * 3.14159265358979323846 / (p->hi[a] - p->lo[a]);
        for (int j = 0; j < p->n[a]; ++j) k[a][j] = dk * ((2 * j < p->n[a]) ? j : j - p->n[a]);
    }
    std::vector<double> t(2 * static_cast<size_t>(p->total));
    for (int kk = 0; kk < p->n[2]; ++kk)
        for (int j = 0; j < p->n[1]; ++j)
            for (int i = 0; i < p->n[0]; ++i) {
                const double kx = k[0][i], ky = k[1][j], kz = k[2][kk];
                const double th = -0.5 * (kx * kx + ky * ky + kz * kz) / p->mass * p->dt;
                const long long f = i + static_cast<long long>(p->n[0]) * (j + static_cast<long long>(p->n[1]) * kk);
                t[2 * f] = std::cos(th);
                t[2 * f + 1] = std::sin(th);
            }
    return h2d(p, p->base_kin, t.data(), t.size() * 8);
}

void release_mask_axes(tdse_plan* p) {
    release(p, p->mask_x, static_cast<uint64_t>(p->n[0]) * p->rsize());
    void* yz = p->mask_yz;
    release(p, yz, static_cast<uint64_t>(p->n[1] + p->n[2]) * 8);
    p->mask_yz = nullptr;
}

// Can the x sweeps use per-axis ramps?  3D and 2D sweep plans whose x lines
// run along the first axis, no magnetic field (its shears take the full
// array).
bool mask_axes_apply(const tdse_plan* p) {
    return (p->kind == kKind3D || (p->kind == kKind2D && p->n[2] == 1)) && p->n[0] > 1 &&
           p->bfield == 0.0 && !(p->flags & TDSE_FLAG_UNFUSED_BASELINE);
}

// A mask equal, bit for bit, to (mx[i] * my[j]) * mz[k] (absorbing_mask,
// core/src/potential.ixx:217-233): the ramps, read off the lines through the
// centre cell (which must be 1), then every cell checked.
bool separable_mask(const tdse_plan* p, const double* m, std::vector<double>* rx,
                    std::vector<double>* ryz) {
    const int n0 = p->n[0], n1 = p->n[1], n2 = p->n[2];
    auto at = [&](int i, int j, int k) {
        return m[i + static_cast<size_t>(n0) * (j + static_cast<size_t>(n1) * k)];
    };
    const int ic = n0 / 2, jc = n1 / 2, kc = n2 / 2;
    if (at(ic, jc, kc) != 1.0) return false;
    rx->assign(static_cast<size_t>(n0), 0.0);
    ryz->assign(static_cast<size_t>(n1 + n2), 0.0);
    for (int i = 0; i < n0; ++i) (*rx)[i] = at(i, jc, kc);
    for (int j = 0; j < n1; ++j) (*ryz)[j] = at(ic, j, kc);
    for (int k = 0; k < n2; ++k) (*ryz)[n1 + k] = at(ic, jc, k);
    for (int k = 0; k < n2; ++k)
        for (int j = 0; j < n1; ++j) {
            const double myz_j = (*ryz)[j], mz_k = (*ryz)[n1 + k];
            for (int i = 0; i < n0; ++i)
                if (at(i, j, k) != ((*rx)[i] * myz_j) * mz_k) return false;
        }
    return true;
}

// Upload the ramps (x in the plan's stored mask form).
int upload_mask_axes(tdse_plan* p, const std::vector<double>& rx, const std::vector<double>& ryz) {
    int rc;
    if (!p->mask_x && (rc = alloc(p, &p->mask_x, static_cast<uint64_t>(p->n[0]) * p->rsize())))
        return rc;
    if (!p->mask_yz) {
        void* yz = nullptr;
        if ((rc = alloc(p, &yz, static_cast<uint64_t>(p->n[1] + p->n[2]) * 8))) return rc;
        p->mask_yz = static_cast<double*>(yz);
    }
    if (p->fp64()) {
        if ((rc = h2d(p, p->mask_x, rx.data(), rx.size() * 8))) return rc;
    } else {
        std::vector<float> f = to_real<float>(rx.data(), static_cast<long long>(rx.size()), kMask);
        if ((rc = h2d(p, p->mask_x, f.data(), f.size() * 4))) return rc;
    }
    return h2d(p, p->mask_yz, ryz.data(), ryz.size() * 8);
}

SweepArgs base_args(const tdse_plan* p) {
    SweepArgs a{};
    a.psi = p->psi;
    a.phi = p->phi;
    a.mask = p->mask;
    a.scale = 1.0 / static_cast<double>(p->total);
    a.psi_bytes = static_cast<uint32_t>(p->total * p->csize());
    a.real_bytes = static_cast<uint32_t>(p->total * p->rsize());
    a.mask_bytes = a.real_bytes;
    for (int i = 0; i < 3; ++i) a.kt[i] = nullptr;
    if (p->imag()) {
        a.relax = p->relax();
        a.partials = p->partials;
    }
    a.cell = 1.0;
    for (int i = 0; i < 3; ++i) a.cell *= (p->hi[i] - p->lo[i]) / p->n[i];
    // the drive's axis . r = sum_a axis_a (lo_a + h_a i_a)
    a.du_c = 0.0;
    double* per[3] = {&a.du_x, &a.du_y, &a.du_z};
    for (int i = 0; i < 3; ++i) {
        a.du_c += p->drive_axis[i] * p->lo[i];
        *per[i] = p->drive_axis[i] * ((p->hi[i] - p->lo[i]) / p->n[i]);
    }
    // x lines are ordered y fastest: line = j + ny k (1D four-step: rows of n1)
    a.dn1 = p->kind == kKindFourStep ? p->n1 : p->n[1];
    return a;
}

double dk_of(const Axis& ax) { return 2.0 * 3.14159265358979323846 / ax.length(); }

// Reduction arguments over the plan's state (x space, empty box).
tdse::ReduceArgs reduce_args(const tdse_plan* p) {
    tdse::ReduceArgs r{};
    r.psi = p->psi;
    r.v = p->has_potential ? p->v : nullptr;
    r.total = p->total;
    for (int a = 0; a < 3; ++a) {
        r.n[a] = p->n[a];
        r.lo[a] = p->lo[a];
        r.h[a] = (p->hi[a] - p->lo[a]) / p->n[a];
        r.sq[a] = a < static_cast<int>(p->sq.size()) ? p->sq[a].n : 1;
        r.dk[a] = a < static_cast<int>(p->sq.size()) ? dk_of(p->sq[a]) : 0.0;
    }
    tdse::empty_box(r);
    r.mass = p->mass;
    return r;
}

double cell_volume(const tdse_plan* p) {
    double c = 1.0;
    for (int a = 0; a < 3; ++a) c *= (p->hi[a] - p->lo[a]) / p->n[a];
    return c;
}


// Sweep descriptors of the plan's geometry.  `target` is the array swept.
struct Geometry {
    SweepArgs x, yf, yi, z;     // 3D / 2D (y carries K in 2D)
    SweepArgs col, row;         // four-step 1D
    SweepArgs single;           // 1D in one block
};

// x sweeps of a plan with per-axis mask ramps read them instead of the array
void use_mask_axes(const tdse_plan* p, SweepArgs* x) {
    if (!p->mask || !p->mask_x || !mask_axes_apply(p)) return;
    x->mask = p->mask_x;
    x->mask_yz = p->mask_yz;
    x->mask_bytes = static_cast<uint32_t>(p->n[0] * p->rsize());
    x->mask_sep = 1;
}

Geometry geometry(const tdse_plan* p, void* target) {
    Geometry g{};
    SweepArgs a = base_args(p);
    a.psi = target;
    if (p->kind == kKind3D) {
        const int n0 = p->sq[0].n, n1 = p->sq[1].n, n2 = p->sq[2].n;
        g.x = a;
        g.x.nlines = static_cast<long long>(n1) * n2;
        g.x.tw = p->tw.at(n0);
        g.yf = a;
        g.yf.nx = n0;
        g.yf.nw = n2;
        g.yf.wstride = static_cast<uint32_t>(n0) * n1;
        g.yf.estride = n0;
        g.yf.tw = p->tw.at(n1);
        g.yi = g.yf;
        g.z = a;
        g.z.nx = n0;
        g.z.nw = n1;
        g.z.wstride = n0;
        g.z.estride = static_cast<uint32_t>(n0) * n1;
        g.z.tw = p->tw.at(n2);
        g.z.kt[0] = p->kt[0];
        g.z.kt[1] = p->kt[1];
        g.z.kt[2] = p->kt[2];
        use_mask_axes(p, &g.x);
    } else if (p->kind == kKind2D) {
        const int n0 = p->sq[0].n, n1 = p->sq[1].n;
        g.x = a;
        g.x.nlines = n1;
        g.x.tw = p->tw.at(n0);
        g.yf = a;
        g.yf.nx = n0;
        g.yf.nw = 1;
        g.yf.wstride = 0;
        g.yf.estride = n0;
        g.yf.tw = p->tw.at(n1);
        g.yf.kt[0] = p->kt[0];
        g.yf.kt[2] = p->kt[1];
        g.yi = g.yf;
        use_mask_axes(p, &g.x);
    } else if (p->kind == kKindSingle) {
        g.single = a;
        g.single.nlines = 1;
        g.single.tw = p->tw.at(p->sq[0].n);
        g.single.kt[2] = p->kt[0];
    } else {
        const int N = p->sq[0].n;
        g.col = a;
        g.col.nx = p->n1;
        g.col.nw = 1;
        g.col.wstride = 0;
        g.col.estride = p->n1;
        g.col.tw = p->tw.at(p->n2);
        g.col.n_total = N;
        g.row = a;
        g.row.nlines = p->n2;
        g.row.tw = p->tw.at(p->n1);
        g.row.kt[2] = p->kt[0];
        g.row.n_total = N;
        g.row.n_rows = p->n2;
    }
    return g;
}

// Sweep of schedule slot `slot` (mode, line length, descriptor); the slots are
// named by kSlots* below.
struct SlotSweep {
    int mode;
    int n;
    SweepArgs args;
};

SlotSweep slot_sweep_real(const tdse_plan* p, const Geometry& g, int slot) {
    if (p->kind == kKind3D) {
        const int n0 = p->sq[0].n, n1 = p->sq[1].n, n2 = p->sq[2].n;
        switch (slot) {
            case 0: return {tdse::C_KICK_FIRST, n0, g.x};
            case 1: return {tdse::S_FWD, n1, g.yf};
            case 2: return {tdse::S_KIN, n2, g.z};
            case 3: return {tdse::S_INV, n1, g.yi};
            case 4: return {tdse::C_KICK_MID, n0, g.x};
            default: return {tdse::C_KICK_LAST, n0, g.x};
        }
    }
    if (p->kind == kKind2D) {
        const int n0 = p->sq[0].n, n1 = p->sq[1].n;
        switch (slot) {
            case 0: return {tdse::C_KICK_FIRST, n0, g.x};
            case 1: return {tdse::S_KIN, n1, g.yf};
            case 2: return {tdse::C_KICK_MID, n0, g.x};
            default: return {tdse::C_KICK_LAST, n0, g.x};
        }
    }
    if (p->kind == kKindSingle) return {tdse::C_SINGLE, p->sq[0].n, g.single};
    switch (slot) {
        case 0: return {tdse::S_KICK_FIRST_TW, p->n2, g.col};
        case 1: return {tdse::C_KIN_TW, p->n1, g.row};
        case 2: return {tdse::S_KICK_MID_TW, p->n2, g.col};
        default: return {tdse::S_KICK_LAST, p->n2, g.col};
    }
}

// Imaginary-time plans run the relax form of the position-space sweeps
// (real weights, a measured norm per step); the other sweeps are shared.
int relax_mode(int mode) {
    switch (mode) {
        case tdse::C_KICK_FIRST: return tdse::C_RELAX_FIRST;
        case tdse::C_KICK_MID: return tdse::C_RELAX_MID;
        case tdse::C_KICK_LAST: return tdse::C_RELAX_LAST;
        case tdse::C_SINGLE: return tdse::C_RELAX_SINGLE;
        case tdse::S_KICK_FIRST_TW: return tdse::S_RELAX_FIRST_TW;
        case tdse::S_KICK_MID_TW: return tdse::S_RELAX_MID_TW;
        case tdse::S_KICK_LAST: return tdse::S_RELAX_LAST;
        default: return mode;
    }
}

SlotSweep slot_sweep(const tdse_plan* p, const Geometry& g, int slot) {
    SlotSweep w = slot_sweep_real(p, g, slot);
    if (p->imag()) w.mode = relax_mode(w.mode);
    return w;
}

// Schedule slots (see kSlots* below) and which of them kick between steps.
int slot_count(const tdse_plan* p) {
    switch (p->kind) {
        case kKind3D: return 6;
        case kKindSingle: return 1;
        case kKindLattice: return 0;
        default: return 4;
    }
}
bool is_kick_slot(const tdse_plan* p, int slot) {
    return p->kind == kKind3D ? slot == 4 : (p->kind != kKindSingle && slot == 2);
}

// The launch list of `nsteps` steps (nsteps >= 1), each sweep in the kernel
// variant chosen for its slot.  Measuring sweeps (imaginary time) carry the
// parity of their step.
std::vector<Launch> schedule(const tdse_plan* p, int nsteps, void* target = nullptr) {
    std::vector<Launch> L;
    const Geometry g = geometry(p, target ? target : p->psi);
    // dipole drive: theta(t) = E0 cos(w t) dt / 2; a call from t0 kicks by
    // -theta(t0) first, -2 theta(t0 + (s + 1) dt) between steps s and s + 1
    // and -theta(t0 + n dt) last (core/src/drive.ixx:52-66)
    const double t0 = p->time;
    auto theta = [&](double t) { return p->drive_amp * std::cos(p->drive_omega * t) * 0.5 * p->dt; };
    auto add = [&](int slot, int step) {
        SlotSweep w = slot_sweep(p, g, slot);
        if (w.mode == tdse::C_SINGLE || w.mode == tdse::C_RELAX_SINGLE) w.args.nsteps = nsteps;
        w.args.parity = step & 1;
        if (p->drive) {
            const bool first = L.empty(), last = slot == slot_count(p) - 1;
            if (w.mode == tdse::C_SINGLE) {
                w.args.damp = p->drive_amp;
                w.args.domega = p->drive_omega;
                w.args.dt0 = t0;
                w.args.ddt = p->dt;
            } else if (first) {
                w.args.dg = -theta(t0);
            } else if (last) {
                w.args.dg = -theta(t0 + nsteps * p->dt);
            } else if (is_kick_slot(p, slot)) {
                w.args.dg = -2.0 * theta(t0 + (step + 1) * p->dt);
            }
        }
        if (p->drive && tdse::mode_kicks(w.mode)) {
            // the driven kick sweeps are compiled once (no variants)
            w.mode |= tdse::kDrive;
            L.push_back({w.mode, w.n, w.args, launcher(p, w.n, w.mode), slot});
        } else {
            L.push_back({w.mode, w.n, w.args, launcher(p, w.n, w.mode, p->variant[slot]), slot});
        }
        // a measuring sweep leaves per-block partials: its norm and the next
        // scale come from one small kernel right after it (timed with its slot)
        if (tdse::mode_measures(w.mode)) L.push_back({kFinishMode, w.n, w.args, &tdse::launch_relax_finish, slot});
    };
    if (p->kind == kKind3D) {
        add(0, 0);
        for (int s = 0; s < nsteps; ++s) {
            add(1, s);
            add(2, s);
            add(3, s);
            add(s + 1 < nsteps ? 4 : 5, s);
        }
    } else if (p->kind == kKindSingle) {
        add(0, 0);
    } else {  // 2D and four-step 1D: first, then (kinetic, mid | last) per step
        add(0, 0);
        for (int s = 0; s < nsteps; ++s) {
            add(1, s);
            add(s + 1 < nsteps ? 2 : 3, s);
        }
    }
    // A finish followed by a sweep that can run it (the next step's first
    // sweep after its measuring x sweep) becomes that sweep's block-0 preamble:
    // one launch fewer per step.  The call's last finish stays a launch.
    std::vector<Launch> F;
    F.reserve(L.size());
    for (size_t i = 0; i < L.size(); ++i) {
        if (L[i].mode == kFinishMode && i + 1 < L.size() && tdse::mode_finishes(L[i + 1].mode)) {
            SweepArgs& next = L[i + 1].args;
            next.fin_partials = L[i].args.partials;
            next.fin_slots = L[i].args.relax;
            next.fin_cell = L[i].args.cell;
            next.fin_parity = L[i].args.parity;
            continue;
        }
        F.push_back(L[i]);
    }
    return F;
}

const char* const kSlots3D[] = {"x_first", "y_fwd", "z_kin", "y_inv", "x_mid", "x_last"};
const char* const kSlots2D[] = {"x_first", "y_kin", "x_mid", "x_last"};
const char* const kSlotsSingle[] = {"line_all_steps"};
const char* const kSlotsFour[] = {"col_first", "row_kin", "col_mid", "col_last"};

int run_launches(tdse_plan* p, const std::vector<Launch>& L) {
    for (const Launch& l : L) {
        if (!l.fn) return fail(TDSE_E_UNSUPPORTED, "no kernel for line length " + std::to_string(l.n));
        TDSE_HIP(l.fn(l.args, p->ctx->max_blocks, p->stream));
    }
    return TDSE_OK;
}

// The reference data flow one pass per operation, FP64
// (core/src/propagator.ixx:110-119, core/src/fft.ixx:102-188): table multiply,
// per-axis forward line FFTs, table multiply, conj, forward FFTs, conj * 1/N,
// table multiply, then the caller's mask.
int run_baseline(tdse_plan* p, int nsteps) {
    Geometry g = geometry(p, p->psi);
    std::vector<Launch> fft;
    auto add = [&](int mode, int n, const SweepArgs& a) {
        fft.push_back({mode, n, a, launcher(p, n, mode), 0});
    };
    add(tdse::C_FWD, p->sq[0].n, g.x);
    add(tdse::S_FWD, p->sq[1].n, g.yf);
    if (p->kind == kKind3D) add(tdse::S_FWD, p->sq[2].n, g.z);
    const double inv = 1.0 / static_cast<double>(p->total);
    for (int s = 0; s < nsteps; ++s) {
        TDSE_HIP(tdse::launch_table_mul(p->psi, p->base_half, p->total, p->stream));
        int rc = run_launches(p, fft);
        if (rc) return rc;
        TDSE_HIP(tdse::launch_table_mul(p->psi, p->base_kin, p->total, p->stream));
        TDSE_HIP(tdse::launch_conj_scale(p->psi, p->total, 1.0, p->stream));
        if ((rc = run_launches(p, fft))) return rc;
        TDSE_HIP(tdse::launch_conj_scale(p->psi, p->total, inv, p->stream));
        TDSE_HIP(tdse::launch_table_mul(p->psi, p->base_half, p->total, p->stream));
        if (p->mask)
            TDSE_HIP(tdse::launch_mask_mul(p->psi, static_cast<const double*>(p->mask), p->total,
                                           p->stream));
    }
    return TDSE_OK;
}

// core/src/magnetic.ixx:34-46: each step is R(a), the static Strang step
// (potential + diamagnetic term), R(a), with a = (B/2)(dt/2); R is the
// three-shear rotation of core/src/rotation.ixx:73-84.  The caller's mask
// follows each whole step, so it rides on the step's last shear.
int run_magnetic(tdse_plan* p, int nsteps) {
    const Geometry g = geometry(p, p->psi);
    const double half_angle = 0.5 * p->bfield * (0.5 * p->dt);
    const int b = (p->b_axis + 1) % 3, c = (p->b_axis + 2) % 3;
    // Shear of the lines along `fa` by coeff * coord along `ca`.
    auto shear = [&](int fa, int ca, double coeff, bool masked) -> Launch {
        SweepArgs a;
        int mode;
        if (fa == 0) {
            a = g.x;
            mode = tdse::C_SHEAR;
            a.sh_sel = ca == 2 ? 1 : 0;  // x lines: y = line % ny, z = line / ny
        } else if (fa == 1) {
            a = g.yf;
            mode = tdse::S_SHEAR;
            a.sh_sel = ca == 2 ? 1 : 0;  // y lines: x = xx, z = w
        } else {
            a = g.z;
            mode = tdse::S_SHEAR;
            a.sh_sel = ca == 1 ? 1 : 0;  // z lines: x = xx, y = w
        }
        const double h = (p->hi[ca] - p->lo[ca]) / p->n[ca];
        a.sh_d0 = coeff * p->lo[ca];
        a.sh_d1 = coeff * h;
        a.dk = 2.0 * 3.14159265358979323846 / (p->hi[fa] - p->lo[fa]);
        a.scale = 1.0 / static_cast<double>(p->n[fa]);
        a.kt[0] = a.kt[1] = a.kt[2] = nullptr;
        a.mask = masked ? p->mask : nullptr;
        const int n = p->n[fa];
        return {mode, n, a, launcher(p, n, mode), 0};
    };
    const double t = std::tan(0.5 * half_angle), sn = std::sin(half_angle);
    std::vector<Launch> L;
    auto rotate = [&](bool mask_last) {
        L.push_back(shear(b, c, -t, false));
        L.push_back(shear(c, b, sn, false));
        L.push_back(shear(b, c, -t, mask_last));
    };
    std::vector<Launch> core = schedule(p, 1);
    core.back().args.mask = nullptr;  // the mask follows the closing rotation
    const bool masked = p->mask != nullptr;
    if (b != 0) {
        for (int s = 0; s < nsteps; ++s) {
            rotate(false);
            for (const Launch& l : core) L.push_back(l);
            rotate(masked);
        }
        return run_launches(p, L);
    }
    // Field along z: the rotation's outer shears run along x, next to the
    // core step's first and last x sweeps, so each joins that sweep (one
    // read and write of the state instead of two); without a mask the
    // closing shear of step s and the opening shear of step s + 1 are one
    // shear by the summed coefficient.
    auto fused = [&](Launch l, int mode, double coeff) {
        const Launch sh = shear(b, c, coeff, false);
        l.mode = mode;
        l.args.sh_sel = sh.args.sh_sel;
        l.args.sh_d0 = sh.args.sh_d0;
        l.args.sh_d1 = sh.args.sh_d1;
        l.args.dk = sh.args.dk;
        l.args.scale = sh.args.scale;
        l.fn = launcher(p, l.n, mode);
        return l;
    };
    const Launch first = fused(core.front(), tdse::C_SHEAR_KICK_FIRST, -t);
    const Launch last = fused(core.back(), tdse::C_KICK_LAST_SHEAR, -t);
    for (int s = 0; s < nsteps; ++s) {
        if (s == 0 || masked) L.push_back(shear(b, c, -t, false));
        L.push_back(shear(c, b, sn, false));
        L.push_back(first);
        for (size_t i = 1; i + 1 < core.size(); ++i) L.push_back(core[i]);
        L.push_back(last);
        L.push_back(shear(c, b, sn, false));
        if (masked)
            L.push_back(shear(b, c, -t, true));
        else
            L.push_back(shear(b, c, s + 1 < nsteps ? -2.0 * t : -t, false));
    }
    return run_launches(p, L);
}

// The Peierls lattice arguments of a plan (core/src/lattice2d.ixx:33-58):
// mixing constants of the bond groups, real time (cos, i sin) or imaginary
// time (cosh, sinh).
tdse::LatticeArgs lattice_args(const tdse_plan* p) {
    tdse::LatticeArgs a{};
    a.psi = p->psi;
    a.phi = p->phi;
    a.mask = p->mask;
    a.v = p->v;
    a.nx = p->n[0];
    a.ny = p->n[1];
    a.tx = p->hop(0);
    a.ty = p->hop(1);
    const double dt = p->dt;
    a.relax = p->imag() ? 1 : 0;
    if (p->imag()) {
        a.cx = std::cosh(a.tx * 0.5 * dt);
        a.mxr = std::sinh(a.tx * 0.5 * dt);
        a.cy = std::cosh(a.ty * 0.5 * dt);
        a.myr = std::sinh(a.ty * 0.5 * dt);
        a.cy2 = std::cosh(a.ty * dt);
        a.my2r = std::sinh(a.ty * dt);
    } else {
        a.cx = std::cos(a.tx * 0.5 * dt);
        a.mxi = std::sin(a.tx * 0.5 * dt);
        a.cy = std::cos(a.ty * 0.5 * dt);
        a.myi = std::sin(a.ty * 0.5 * dt);
        a.cy2 = std::cos(a.ty * dt);
        a.my2i = std::sin(a.ty * dt);
    }
    a.link_mode = p->lat_mode;
    a.link_a = p->lat_a;
    a.ylo = p->lo[1];
    a.hy = (p->hi[1] - p->lo[1]) / p->n[1];
    a.link_col = p->lat_col;
    a.link_j0 = p->lat_j0;
    a.link_j1 = p->lat_j1;
    if (p->lat_mode == 2) {  // the cut's link exp(lat_sign i phi), as std::polar forms it
        a.link_c = std::cos(p->lat_a);
        a.link_s = p->lat_sign * std::sin(p->lat_a);
    }
    return a;
}

// core/src/lattice2d.ixx:218-249.  A call opens with the half step P X0 X1
// (one row sweep), then runs whole steps on 2D tiles, two per launch, out of
// place between psi and a scratch array; the last step ends with its P and
// the mask.  Imaginary time: each step is a call of one step, renormalised
// after it (fixed-order reduction, then a scaling pass).
int run_lattice(tdse_plan* p, int nsteps) {
    int rc;
    if (!p->scratch && (rc = alloc(p, &p->scratch, p->total * p->csize()))) return rc;
    tdse::LatticeArgs a = lattice_args(p);
    const bool fp64 = p->fp64();
    // the opening row sweep writes where the tile launches must start so that
    // the last one writes psi
    auto steps = [&](int n) {
        const int per = tdse::kLatStepsPerLaunch;
        const int launches = (n + per - 1) / per;
        void* cur = (launches & 1) ? p->scratch : p->psi;
        TDSE_HIP(tdse::launch_lattice_rows(fp64, a, p->psi, cur, p->stream));
        for (int l = 0; l < launches; ++l) {
            void* next = cur == p->psi ? p->scratch : p->psi;
            const int units = std::min(per, n - l * per);
            TDSE_HIP(tdse::launch_lattice_steps(fp64, a, cur, next, units,
                                                l + 1 < launches ? 1 : 2, p->stream));
            cur = next;
        }
        return TDSE_OK;
    };
    if (!p->imag()) return steps(nsteps);
    const tdse::ReduceArgs r = reduce_args(p);
    const double dv = cell_volume(p);
    for (int s = 0; s < nsteps; ++s) {
        if ((rc = steps(1))) return rc;
        TDSE_HIP(tdse::launch_reduce(fp64, false, r, p->partials, p->red_out, p->stream));
        TDSE_HIP(tdse::launch_normalize_dev(fp64, p->psi, p->total, p->red_out, dv,
                                            &p->relax()->norm_sq, p->stream));
    }
    return TDSE_OK;
}

int check_plan(const tdse_plan* p) {
    if (!p) return fail(TDSE_E_INVALID, "null plan");
    return TDSE_OK;
}

}  // namespace

// ---- batches of plans ---------------------------------------------------------------

namespace {

// Blocks per batched launch; each block sweeps a contiguous chunk of tiles
// (one tile per block up to this many tiles: the seven 160^3 FP32 members of
// the uneven batch sweep in 0.478 ms per round with one tile per block
// against 0.513 ms with chunks of 2048 blocks).
#ifndef TDSE_BATCH_BLOCKS
#define TDSE_BATCH_BLOCKS 65536
#endif
constexpr int kBatchMaxBlocks = TDSE_BATCH_BLOCKS;

// Streams a batch spreads its launches over (the main one included).
#ifndef TDSE_BATCH_LANES
#define TDSE_BATCH_LANES 4
#endif
// Working-set cap of a unit (bytes of state, kick phase and mask that one
// step touches): a quarter of the 256 MB Infinity Cache, one lane's share.
#ifndef TDSE_BATCH_UNIT_BYTES
#define TDSE_BATCH_UNIT_BYTES (64LL << 20)
#endif

// Plans that share (precision, line length, mode) in one sub-phase of a round
// are swept by one batched launch.
struct BatchGroup {
    int precision = 0, mode = 0, n = 0, variant = 0;
    tdse::BatchLaunchFn fn = nullptr;
    std::vector<int> plans;       // indices into the batch's plan list
    void* d_items = nullptr;  // SweepArgs[nitems]
    void* d_first = nullptr;  // long long[nitems + 1]
    int nitems = 0;
    long long tiles = 0;
    tdse::BatchTable table(int nsteps) const {
        return {static_cast<const SweepArgs*>(d_items), static_cast<const long long*>(d_first),
                nitems, nsteps};
    }
};

// One sweep of a plan inside a round: sub-phase 0..3, modes for an interior
// round and for the last round of a call, and the plan's kernel variants of
// those sweeps (the schedule slots they stand for).
struct RoundSweep {
    int sub, n, mode_mid, mode_last;
    SweepArgs args;
    int var_mid, var_last;
};

std::vector<RoundSweep> plan_round(const tdse_plan* p) {
    Geometry g = geometry(p, p->psi);
    const int* v = p->variant;
    std::vector<RoundSweep> r;
    if (p->kind == kKind3D) {
        r.push_back({0, p->sq[1].n, tdse::S_FWD, tdse::S_FWD, g.yf, v[1], v[1]});
        r.push_back({1, p->sq[2].n, tdse::S_KIN, tdse::S_KIN, g.z, v[2], v[2]});
        r.push_back({2, p->sq[1].n, tdse::S_INV, tdse::S_INV, g.yi, v[3], v[3]});
        r.push_back({3, p->sq[0].n, tdse::C_KICK_MID, tdse::C_KICK_LAST, g.x, v[4], v[5]});
    } else if (p->kind == kKind2D) {
        r.push_back({1, p->sq[1].n, tdse::S_KIN, tdse::S_KIN, g.yf, v[1], v[1]});
        r.push_back({3, p->sq[0].n, tdse::C_KICK_MID, tdse::C_KICK_LAST, g.x, v[2], v[3]});
    } else if (p->kind == kKindFourStep) {
        r.push_back({1, p->n1, tdse::C_KIN_TW, tdse::C_KIN_TW, g.row, v[1], v[1]});
        r.push_back({3, p->n2, tdse::S_KICK_MID_TW, tdse::S_KICK_LAST, g.col, v[2], v[3]});
    }
    return r;
}

bool plan_init(const tdse_plan* p, int* n, int* mode, SweepArgs* a, int* variant) {
    Geometry g = geometry(p, p->psi);
    *variant = p->variant[0];
    if (p->kind == kKind3D || p->kind == kKind2D) {
        *n = p->sq[0].n;
        *mode = tdse::C_KICK_FIRST;
        *a = g.x;
        return true;
    }
    if (p->kind == kKindFourStep) {
        *n = p->n2;
        *mode = tdse::S_KICK_FIRST_TW;
        *a = g.col;
        return true;
    }
    return false;
}

long long tiles_of(int precision, int n, int mode, const SweepArgs& a, int variant = 0) {
    const tdse::SweepShape sh = tdse::sweep_shape(precision == TDSE_FP64, n, mode, variant);
    if (sh.per_block <= 0) return 0;
    if (tdse::mode_is_contiguous(mode)) return (a.nlines + sh.per_block - 1) / sh.per_block;
    return static_cast<long long>((a.nx + sh.per_block - 1) / sh.per_block) * a.nw;
}

}  // namespace

// A unit of a batch: members that advance together, a call's steps at a
// time, on one lane.  Its groups sweep all of its members per launch.
struct BatchUnit {
    std::vector<int> members;  // indices into the batch's plan list
    std::vector<BatchGroup> init, mid[4], last[4];
    int lane = 0;
    uint64_t bytes = 0;        // working set of one step (state, kick phase, mask)
};

struct tdse_batch {
    tdse_ctx* ctx = nullptr;
    std::vector<tdse_plan*> plans;
    hipStream_t stream = nullptr;
    static constexpr int kSide = TDSE_BATCH_LANES - 1;
    hipStream_t side[kSide] = {};
    hipEvent_t fork = nullptr;
    hipEvent_t done = nullptr;  // end of the last call, awaited by the plans' streams
    hipEvent_t join[kSide] = {};
    std::vector<BatchUnit> units;
    uint64_t bytes = 0;
};

namespace {

int batch_alloc(tdse_batch* b, void** out, uint64_t bytes) {
    if (hipMalloc(out, bytes) != hipSuccess) {
        (void)hipGetLastError();
        *out = nullptr;
        return fail(TDSE_E_OOM, "hipMalloc for batch tables failed");
    }
    b->bytes += bytes;
    return TDSE_OK;
}

int upload_group(tdse_batch* b, BatchGroup& g, const std::vector<SweepArgs>& items) {
    g.nitems = static_cast<int>(items.size());
    std::vector<long long> first(items.size() + 1, 0);
    for (size_t i = 0; i < items.size(); ++i)
        first[i + 1] = first[i] + tiles_of(g.precision, g.n, g.mode, items[i], g.variant);
    g.tiles = first.back();
    int rc;
    if ((rc = batch_alloc(b, &g.d_items, items.size() * sizeof(SweepArgs)))) return rc;
    if ((rc = batch_alloc(b, &g.d_first, first.size() * sizeof(long long)))) return rc;
    // ordered on the batch stream (non-blocking, like the lanes that wait on it)
    TDSE_HIP(hipMemcpyAsync(g.d_items, items.data(), items.size() * sizeof(SweepArgs),
                            hipMemcpyHostToDevice, b->stream));
    TDSE_HIP(hipMemcpyAsync(g.d_first, first.data(), first.size() * sizeof(long long),
                            hipMemcpyHostToDevice, b->stream));
    TDSE_HIP(hipStreamSynchronize(b->stream));
    g.fn = tdse::find_batch_launcher(g.precision == TDSE_FP64, g.n, g.mode, g.variant);
    if (!g.fn) return fail(TDSE_E_UNSUPPORTED, "no batched kernel for line length " + std::to_string(g.n));
    return TDSE_OK;
}

// One plan's sweep in a sub-phase: line length, mode, kernel variant, descriptor.
struct BatchRow {
    const tdse_plan* plan;
    int n, mode, variant;
    SweepArgs args;
};

// Group (precision, n, mode, variant) -> items, in first-seen order
// (deterministic).  A variant the batched kernels lack falls back to 0.
int build_groups(tdse_batch* b, std::vector<BatchGroup>& out, const std::vector<BatchRow>& rows) {
    std::vector<BatchGroup> groups;
    std::vector<std::vector<SweepArgs>> items;
    for (const BatchRow& row : rows) {
        const int prec = row.plan->precision, n = row.n, mode = row.mode;
        const int var = tdse::find_batch_launcher(prec == TDSE_FP64, n, mode, row.variant)
                            ? row.variant
                            : 0;
        size_t gi = 0;
        for (; gi < groups.size(); ++gi)
            if (groups[gi].precision == prec && groups[gi].n == n && groups[gi].mode == mode &&
                groups[gi].variant == var)
                break;
        if (gi == groups.size()) {
            BatchGroup g;
            g.precision = prec;
            g.n = n;
            g.mode = mode;
            g.variant = var;
            groups.push_back(g);
            items.emplace_back();
        }
        items[gi].push_back(row.args);
        const auto it = std::find(b->plans.begin(), b->plans.end(), row.plan);
        groups[gi].plans.push_back(static_cast<int>(it - b->plans.begin()));
    }
    for (size_t gi = 0; gi < groups.size(); ++gi) {
        int rc = upload_group(b, groups[gi], items[gi]);
        if (rc) return rc;
        out.push_back(groups[gi]);
    }
    return TDSE_OK;
}

hipStream_t lane_stream(const tdse_batch* b, int lane) {
    return lane == 0 ? b->stream : b->side[lane - 1];
}

// Launch groups one after another on `st` (stream order is the dependency:
// a unit's sweeps of one sub-phase never touch the same plan twice, and the
// next sub-phase follows them on the same stream).
int launch_groups(const std::vector<BatchGroup>& gs, hipStream_t st, int nsteps = 1) {
    for (const BatchGroup& g : gs) TDSE_HIP(g.fn(g.table(nsteps), g.tiles, kBatchMaxBlocks, st));
    return TDSE_OK;
}

// A call of `nsteps` steps: each lane runs its units one after another, each
// unit's opening sweeps, interior rounds and closing round.  The host queues
// the lanes' work in slices of a few rounds, lane by lane, so every lane has
// work queued from the start of the call.
int launch_units(const tdse_batch* b, int nsteps) {
    constexpr int kSliceRounds = 8;
    const int lanes = tdse_batch::kSide + 1;
    std::vector<std::vector<const BatchUnit*>> per_lane(static_cast<size_t>(lanes));
    for (const BatchUnit& u : b->units) per_lane[static_cast<size_t>(u.lane)].push_back(&u);
    std::vector<size_t> unit(static_cast<size_t>(lanes), 0);
    std::vector<int> round(static_cast<size_t>(lanes), -1);  // -1: opening sweeps not queued
    for (bool queued = true; queued;) {
        queued = false;
        for (int l = 0; l < lanes; ++l) {
            if (unit[l] >= per_lane[l].size()) continue;
            queued = true;
            const BatchUnit& u = *per_lane[l][unit[l]];
            const hipStream_t st = lane_stream(b, l);
            int rc;
            if (round[l] < 0) {
                // the one-block 1D groups take nsteps and run the whole call
                if ((rc = launch_groups(u.init, st, nsteps))) return rc;
                round[l] = 0;
            }
            const int end = std::min(nsteps, round[l] + kSliceRounds);
            for (int r = round[l]; r < end; ++r) {
                const bool last = r + 1 == nsteps;
                for (int sub = 0; sub < 4; ++sub)
                    if ((rc = launch_groups(last ? u.last[sub] : u.mid[sub], st))) return rc;
            }
            round[l] = end;
            if (end == nsteps) {
                ++unit[l];
                round[l] = -1;
            }
        }
    }
    return TDSE_OK;
}

// Working set of one step of a plan: state, kick phase, mask array.
uint64_t step_working_set(const tdse_plan* p) {
    uint64_t b = static_cast<uint64_t>(p->total) * (p->csize() + p->rsize());
    if (p->mask && !(p->mask_x && mask_axes_apply(p))) b += static_cast<uint64_t>(p->total) * p->rsize();
    return b;
}

// Same unit class: members that can share launches (precision, geometry kind,
// line lengths).
bool same_class(const tdse_plan* a, const tdse_plan* b) {
    if (a->precision != b->precision || a->kind != b->kind || a->sq.size() != b->sq.size())
        return false;
    for (size_t i = 0; i < a->sq.size(); ++i)
        if (a->sq[i].n != b->sq[i].n) return false;
    return a->n1 == b->n1 && a->n2 == b->n2;
}

}  // namespace

extern "C" {

const char* tdse_version(void) { return "tdse-mi355x 0.1 (gfx950)"; }

const char* tdse_status_string(int status) {
    switch (status) {
        case TDSE_OK: return "ok";
        case TDSE_E_INVALID: return "invalid argument";
        case TDSE_E_UNSUPPORTED: return "unsupported shape";
        case TDSE_E_OOM: return "out of device memory or budget";
        case TDSE_E_DEVICE: return "device error";
        case TDSE_E_TIMEOUT: return "timeout";
        case TDSE_E_NO_DEVICE: return "no such device";
        case TDSE_E_STATE: return "call out of order";
        default: return "unknown status";
    }
}

const char* tdse_last_error(void) { return g_last_error.c_str(); }

int tdse_ctx_create(int device_ordinal, uint64_t hbm_budget_bytes, tdse_ctx** out) {
    if (!out) return fail(TDSE_E_INVALID, "null out");
    *out = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || device_ordinal < 0 || device_ordinal >= count) {
        (void)hipGetLastError();
        return fail(TDSE_E_NO_DEVICE, "no HIP device " + std::to_string(device_ordinal));
    }
    tdse_ctx* c = new tdse_ctx();
    c->device = device_ordinal;
    c->budget = hbm_budget_bytes;
    *out = c;
    return TDSE_OK;
}

int tdse_ctx_destroy(tdse_ctx* ctx) {
    if (!ctx) return fail(TDSE_E_INVALID, "null ctx");
    (void)hipSetDevice(ctx->device);
    for (int b = 0; b < 2; ++b) {
        if (ctx->pin.buf[b]) (void)hipHostFree(ctx->pin.buf[b]);
        if (ctx->pin.ev[b]) (void)hipEventDestroy(ctx->pin.ev[b]);
    }
    delete ctx;
    return TDSE_OK;
}

int tdse_ctx_bytes_in_use(const tdse_ctx* ctx, uint64_t* out) {
    if (!ctx || !out) return fail(TDSE_E_INVALID, "null argument");
    std::lock_guard<std::mutex> lk(const_cast<tdse_ctx*>(ctx)->mu);
    *out = ctx->used;
    return TDSE_OK;
}

int tdse_plan_bytes(const int n[3], int precision, int with_mask, uint64_t* out) {
    if (!n || !out) return fail(TDSE_E_INVALID, "null argument");
    if (precision != TDSE_FP32 && precision != TDSE_FP64) return fail(TDSE_E_INVALID, "precision");
    std::vector<Axis> sq;
    Kind kind;
    int n1 = 0, n2 = 0;
    std::vector<int> lines;
    for (int a = 0; a < 3; ++a)
        if (n[a] < 1) return fail(TDSE_E_INVALID, "n < 1");
    int rc = classify(n, sq, nullptr, nullptr, &kind, &n1, &n2, &lines);
    if (rc) return rc;
    const long long total = static_cast<long long>(n[0]) * n[1] * n[2];
    long long ktp = 0;
    for (const Axis& ax : sq) ktp += ax.n;
    // the x sweeps' mask ramps (mask_axes_apply of a plan of this shape)
    const bool ramps = (kind == kKind3D || (kind == kKind2D && n[2] == 1)) && n[0] > 1;
    *out = plan_bytes_for(n, total, lines, precision == TDSE_FP64, with_mask != 0, ktp, ramps);
    return TDSE_OK;
}

int tdse_plan_create(tdse_ctx* ctx, const int n[3], const double lo[3], const double hi[3],
                     double dt, double mass, int precision, unsigned flags, tdse_plan** out) {
    if (!ctx || !n || !lo || !hi || !out) return fail(TDSE_E_INVALID, "null argument");
    if (flags & ~(TDSE_FLAG_UNFUSED_BASELINE | TDSE_FLAG_IMAGINARY_TIME | TDSE_FLAG_LATTICE))
        return fail(TDSE_E_INVALID, "unknown flags");
    if (flags & TDSE_FLAG_LATTICE) {
        if ((flags & TDSE_FLAG_UNFUSED_BASELINE) || !n)
            return fail(TDSE_E_INVALID, "the lattice has no unfused baseline");
        if (n[2] != 1 || n[0] < 2 || n[1] < 2)
            return fail(TDSE_E_INVALID, "the lattice is a 2D grid n x m x 1 with n, m >= 2");
        if (n[0] > tdse::kLatMaxNx) return fail(TDSE_E_UNSUPPORTED, "lattice rows hold at most 2048 sites");
    }
    if ((flags & TDSE_FLAG_UNFUSED_BASELINE) && precision != TDSE_FP64)
        return fail(TDSE_E_INVALID, "the unfused baseline is FP64 only");
    if ((flags & TDSE_FLAG_UNFUSED_BASELINE) && (flags & TDSE_FLAG_IMAGINARY_TIME))
        return fail(TDSE_E_INVALID, "the unfused baseline is real time only");
    *out = nullptr;
    if (precision != TDSE_FP32 && precision != TDSE_FP64)
        return fail(TDSE_E_INVALID, "precision must be TDSE_FP32 or TDSE_FP64");
    if (!std::isfinite(dt) || !std::isfinite(mass) || mass == 0.0)
        return fail(TDSE_E_INVALID, "dt and mass must be finite, mass non-zero");
    for (int a = 0; a < 3; ++a) {
        if (n[a] < 1) return fail(TDSE_E_INVALID, "grid extent < 1");
        if (!(hi[a] > lo[a]) && n[a] > 1) return fail(TDSE_E_INVALID, "hi must exceed lo");
    }
    const long long total = static_cast<long long>(n[0]) * n[1] * n[2];
    if (total * (precision == TDSE_FP64 ? 16 : 8) >= (1LL << 32))
        return fail(TDSE_E_UNSUPPORTED, "state larger than 4 GiB (32-bit buffer offsets)");
    int rc = set_device(ctx);
    if (rc) return rc;

    tdse_plan* p = new tdse_plan();
    p->ctx = ctx;
    p->precision = precision;
    for (int a = 0; a < 3; ++a) {
        p->n[a] = n[a];
        p->lo[a] = lo[a];
        p->hi[a] = hi[a];
    }
    p->dt = dt;
    p->mass = mass;
    p->total = total;
    p->flags = flags;
    std::vector<int> lines;
    if (flags & TDSE_FLAG_LATTICE) {
        // finite differences: any extents, no line FFTs
        p->kind = kKindLattice;
        for (int a = 0; a < 2; ++a) p->sq.push_back({n[a], lo[a], hi[a]});
    } else {
        rc = classify(n, p->sq, lo, hi, &p->kind, &p->n1, &p->n2, &lines);
        if (rc) {
            delete p;
            return rc;
        }
    }
    if ((flags & TDSE_FLAG_UNFUSED_BASELINE) && p->kind != kKind3D && p->kind != kKind2D) {
        delete p;
        return fail(TDSE_E_UNSUPPORTED, "the unfused baseline covers 2D and 3D grids");
    }
    auto bail = [&](int code) {
        tdse_plan_destroy(p);
        return code;
    };
    if (hipStreamCreateWithFlags(&p->stream, hipStreamNonBlocking) != hipSuccess) {
        (void)hipGetLastError();
        delete p;
        return fail(TDSE_E_DEVICE, "hipStreamCreate failed");
    }
    register_stream(ctx, p->stream);
    if ((rc = alloc(p, &p->psi, total * p->csize()))) return bail(rc);
    if ((rc = alloc(p, &p->phi, total * p->rsize()))) return bail(rc);
    if ((rc = alloc(p, &p->v, total * p->rsize()))) return bail(rc);
    void* part = nullptr;
    if ((rc = alloc(p, &part, static_cast<uint64_t>(tdse::kMaxReduceBlocks) * tdse::kReduceQ * 8 +
                                  64 * 8)))
        return bail(rc);
    p->partials = static_cast<double*>(part);
    p->red_out = p->partials + tdse::kMaxReduceBlocks * tdse::kReduceQ;
    static_assert(2 * tdse::kRelaxMaxBlocks <= tdse::kMaxReduceBlocks * tdse::kReduceQ,
                  "relax partials share the reduction partials");
    static_assert(16 * 8 + sizeof(tdse::RelaxSlots) <= 64 * 8, "relax slots follow red_out");
    for (int N : lines)
        if ((rc = make_twiddles(p, N))) return bail(rc);
    if (p->kind != kKindLattice)
        for (size_t a = 0; a < p->sq.size(); ++a)
            if ((rc = make_kinetic(p, static_cast<int>(a), p->sq[a]))) return bail(rc);
    if (flags & TDSE_FLAG_UNFUSED_BASELINE) {
        if ((rc = alloc(p, &p->base_half, total * 16))) return bail(rc);
        if ((rc = alloc(p, &p->base_kin, total * 16))) return bail(rc);
        if ((rc = make_baseline_kinetic(p))) return bail(rc);
    }
    if (hipMemsetAsync(p->psi, 0, total * p->csize(), p->stream) != hipSuccess ||
        hipMemsetAsync(p->phi, 0, total * p->rsize(), p->stream) != hipSuccess ||
        hipMemsetAsync(p->v, 0, total * p->rsize(), p->stream) != hipSuccess ||
        hipMemsetAsync(p->partials, 0,
                       static_cast<uint64_t>(tdse::kMaxReduceBlocks) * tdse::kReduceQ * 8 + 64 * 8,
                       p->stream) != hipSuccess ||
        hipStreamSynchronize(p->stream) != hipSuccess) {
        (void)hipGetLastError();
        return bail(fail(TDSE_E_DEVICE, "hipMemset failed"));
    }
    *out = p;
    return TDSE_OK;
}

int tdse_plan_destroy(tdse_plan* p) {
    if (!p) return fail(TDSE_E_INVALID, "null plan");
    (void)hipSetDevice(p->ctx->device);
    if (p->stream) (void)hipStreamSynchronize(p->stream);
    const uint64_t cs = p->csize(), rs = p->rsize();
    release(p, p->psi, p->total * cs);
    release(p, p->phi, p->total * rs);
    release(p, p->v, p->total * rs);
    release(p, p->mask, p->total * rs);
    release_mask_axes(p);
    release(p, p->scratch, p->total * cs);
    release(p, p->kvar, p->total * 16);
    release(p, p->base_half, p->total * 16);
    release(p, p->base_kin, p->total * 16);
    void* part = p->partials;
    release(p, part, static_cast<uint64_t>(tdse::kMaxReduceBlocks) * tdse::kReduceQ * 8 + 64 * 8);
    for (auto& kv : p->tw) release(p, kv.second, static_cast<uint64_t>(kv.first) * 16);
    for (int a = 0; a < 3; ++a) release(p, p->kt[a], static_cast<uint64_t>(p->kt_len[a]) * 16);
    if (p->stream) {
        unregister_stream(p->ctx, p->stream);
        (void)hipStreamDestroy(p->stream);
    }
    delete p;
    return TDSE_OK;
}

int tdse_plan_set_potential(tdse_plan* p, const double* v) {
    int rc = check_plan(p);
    if (rc) return rc;
    if (!v) return fail(TDSE_E_INVALID, "null potential");
    if ((rc = set_device(p->ctx))) return rc;
    TDSE_HIP(hipStreamSynchronize(p->stream));
    std::vector<double> veff;
    if (p->bfield != 0.0) {
        // the diamagnetic term joins the potential (core/src/magnetic.ixx:50-73)
        veff.assign(v, v + p->total);
        const double c = p->bfield * p->bfield / 8.0;
        for (int k = 0; k < p->n[2]; ++k)
            for (int j = 0; j < p->n[1]; ++j)
                for (int i = 0; i < p->n[0]; ++i) {
                    const int idx[3] = {i, j, k};
                    double perp2 = 0.0;
                    for (int a = 0; a < 3; ++a) {
                        if (a == p->b_axis) continue;
                        const double x = p->lo[a] + idx[a] * ((p->hi[a] - p->lo[a]) / p->n[a]);
                        perp2 += x * x;
                    }
                    veff[i + static_cast<size_t>(p->n[0]) * (j + static_cast<size_t>(p->n[1]) * k)] +=
                        c * perp2;
                }
        v = veff.data();
    }
    std::vector<double> th(static_cast<size_t>(p->total));
    // the lattice onsite energy carries the finite-difference diagonal
    // 2 tx + 2 ty (core/src/lattice2d.ixx:50-56)
    const double shift = p->kind == kKindLattice ? 2.0 * p->hop(0) + 2.0 * p->hop(1) : 0.0;
    for (long long i = 0; i < p->total; ++i) {
        if (!std::isfinite(v[i])) return fail(TDSE_E_INVALID, "potential not finite");
        th[i] = -0.5 * (v[i] + shift) * p->dt;  // core/src/propagator.ixx:29, 79
    }
    // real time: reduced modulo 2 pi in FP64 for both precisions (the FP64
    // kernels' sin/cos take |angle| <= 8), FP32 sweep plans then keep its
    // half-angle tangent; imaginary time: the exponent of the weight
    // exp(-V dtau / 2) (core/src/imaginary_time.ixx:30-33, 75-78)
    if ((rc = upload_real(p, p->phi, th.data(), p->imag() ? kPlain : p->tan_kick() ? kTanHalf : kPhase)))
        return rc;
    if ((rc = upload_real(p, p->v, v, kPlain))) return rc;
    if (p->base_half) {
        std::vector<double> h(2 * static_cast<size_t>(p->total));
        for (long long i = 0; i < p->total; ++i) {
            h[2 * i] = std::cos(th[i]);
            h[2 * i + 1] = std::sin(th[i]);
        }
        if ((rc = h2d(p, p->base_half, h.data(), h.size() * 8))) return rc;
    }
    p->has_potential = true;
    return TDSE_OK;
}

int tdse_plan_set_mask(tdse_plan* p, const double* mask) {
    int rc = check_plan(p);
    if (rc) return rc;
    if (mask && p->imag()) return fail(TDSE_E_INVALID, "imaginary-time plans take no mask");
    if ((rc = set_device(p->ctx))) return rc;
    TDSE_HIP(hipStreamSynchronize(p->stream));
    release_mask_axes(p);
    if (!mask) {
        release(p, p->mask, p->total * p->rsize());
        return TDSE_OK;
    }
    if (!p->mask && (rc = alloc(p, &p->mask, p->total * p->rsize()))) return rc;
    if ((rc = upload_real(p, p->mask, mask, kMask))) return rc;
    std::vector<double> rx, ryz;
    if (mask_axes_apply(p) && separable_mask(p, mask, &rx, &ryz)) return upload_mask_axes(p, rx, ryz);
    return TDSE_OK;
}

int tdse_plan_build(tdse_plan* p, const tdse_recipe* r) {
    int rc = check_plan(p);
    if (rc) return rc;
    if (!r) return fail(TDSE_E_INVALID, "null recipe");
    if (p->flags & TDSE_FLAG_UNFUSED_BASELINE)
        return fail(TDSE_E_INVALID, "the unfused baseline takes host inputs");
    if (p->imag() && r->mask_width > 0.0)
        return fail(TDSE_E_INVALID, "imaginary-time plans take no mask");
    for (int a = 0; a < 3; ++a)
        if (!(r->sigma[a] > 0.0)) return fail(TDSE_E_INVALID, "sigma must be positive");
    if ((rc = set_device(p->ctx))) return rc;
    TDSE_HIP(hipStreamSynchronize(p->stream));
    release_mask_axes(p);
    if (r->mask_width > 0.0) {
        if (!p->mask && (rc = alloc(p, &p->mask, p->total * p->rsize()))) return rc;
        if (mask_axes_apply(p)) {
            // the ramps of absorbing_mask (core/src/potential.ixx:194-215)
            // from the host builder, one axis at a time
            std::vector<double> rx(static_cast<size_t>(p->n[0])), ryz(static_cast<size_t>(p->n[1] + p->n[2]));
            for (int a = 0; a < 3; ++a) {
                const int n1d[3] = {p->n[a], 1, 1};
                const double lo1d[3] = {p->lo[a], 0.0, 0.0}, hi1d[3] = {p->hi[a], 1.0, 1.0};
                double* dst = a == 0 ? rx.data() : ryz.data() + (a == 1 ? 0 : p->n[1]);
                if (tdse_build_absorbing_mask(n1d, lo1d, hi1d, r->mask_width, dst) != TDSE_OK)
                    return fail(TDSE_E_INVALID, "bad mask width");
            }
            if ((rc = upload_mask_axes(p, rx, ryz))) return rc;
        }
    } else {
        release(p, p->mask, p->total * p->rsize());
    }
    tdse::BuildArgs b{};
    b.phi = p->phi;
    b.v = p->v;
    b.mask = p->mask;
    b.psi = p->psi;
    b.total = p->total;
    for (int a = 0; a < 3; ++a) {
        b.n[a] = p->n[a];
        b.lo[a] = p->lo[a];
        b.hi[a] = p->hi[a];
        b.h[a] = (p->hi[a] - p->lo[a]) / p->n[a];
    }
    b.dt = p->dt;
    b.imaginary = p->imag() ? 1 : 0;
    b.tan_half = p->tan_kick() ? 1 : 0;
    b.bfield = p->bfield;
    b.b_axis = p->b_axis;
    b.onsite_shift = p->kind == kKindLattice ? 2.0 * p->hop(0) + 2.0 * p->hop(1) : 0.0;
    b.r = *r;
    TDSE_HIP(tdse::launch_build(p->fp64(), b, p->stream));
    p->has_potential = true;
    // normalise the packet: sum |psi|^2 dV -> scale (core/src/field.ixx:78-87)
    tdse_observables o;
    if ((rc = tdse_observe(p, TDSE_OBS_NORM, &o))) return rc;
    if (o.norm > 0.0) TDSE_HIP(tdse::launch_scale(p->fp64(), p->psi, p->total, 1.0 / std::sqrt(o.norm), p->stream));
    return TDSE_OK;
}

int tdse_state_upload(tdse_plan* p, const double* psi) {
    int rc = check_plan(p);
    if (rc) return rc;
    if (!psi) return fail(TDSE_E_INVALID, "null state");
    if ((rc = set_device(p->ctx))) return rc;
    {
        std::lock_guard<std::mutex> lk(p->ctx->pin.mu);
        if (pinned_stage(p->ctx))
            return p->fp64() ? upload_staged<double>(p, p->psi, psi, 2 * p->total)
                             : upload_staged<float>(p, p->psi, psi, 2 * p->total);
    }
    if (p->fp64()) return h2d(p, p->psi, psi, p->total * 16);
    // left uninitialised: the converting threads touch its pages first
    std::unique_ptr<float[]> f(new float[2 * static_cast<size_t>(p->total)]);
    parallel_chunks(2 * p->total, [&](long long i0, long long i1) {
        for (long long i = i0; i < i1; ++i) f[i] = static_cast<float>(psi[i]);
    });
    return h2d(p, p->psi, f.get(), 2 * static_cast<size_t>(p->total) * 4);
}

int tdse_state_download(tdse_plan* p, double* psi) {
    int rc = check_plan(p);
    if (rc) return rc;
    if (!psi) return fail(TDSE_E_INVALID, "null state");
    if ((rc = set_device(p->ctx))) return rc;
    {
        std::lock_guard<std::mutex> lk(p->ctx->pin.mu);
        if (pinned_stage(p->ctx))
            return p->fp64() ? download_staged<double>(p, psi, p->psi, 2 * p->total)
                             : download_staged<float>(p, psi, p->psi, 2 * p->total);
    }
    if (p->fp64()) return d2h(p, psi, p->psi, p->total * 16);
    std::unique_ptr<float[]> f(new float[2 * static_cast<size_t>(p->total)]);
    if ((rc = d2h(p, f.get(), p->psi, 2 * static_cast<size_t>(p->total) * 4))) return rc;
    parallel_chunks(2 * p->total, [&](long long i0, long long i1) {
        for (long long i = i0; i < i1; ++i) psi[i] = f[i];
    });
    return TDSE_OK;
}

int tdse_state_download_f32(tdse_plan* p, float* psi) {
    int rc = check_plan(p);
    if (rc) return rc;
    if (!psi) return fail(TDSE_E_INVALID, "null state");
    if ((rc = set_device(p->ctx))) return rc;
    if (!p->fp64()) return d2h(p, psi, p->psi, 2 * static_cast<size_t>(p->total) * 4);
    {
        std::lock_guard<std::mutex> lk(p->ctx->pin.mu);
        if (pinned_stage(p->ctx)) return download_staged<double, float>(p, psi, p->psi, 2 * p->total);
    }
    std::unique_ptr<double[]> d(new double[2 * static_cast<size_t>(p->total)]);
    if ((rc = d2h(p, d.get(), p->psi, 2 * static_cast<size_t>(p->total) * 8))) return rc;
    parallel_chunks(2 * p->total, [&](long long i0, long long i1) {
        for (long long i = i0; i < i1; ++i) psi[i] = static_cast<float>(d[i]);
    });
    return TDSE_OK;
}

int tdse_state_copy(tdse_plan* dst, const tdse_plan* src) {
    if (!dst || !src) return fail(TDSE_E_INVALID, "null plan");
    if (dst->total != src->total || dst->precision != src->precision)
        return fail(TDSE_E_INVALID, "shape or precision mismatch");
    int rc = set_device(dst->ctx);
    if (rc) return rc;
    TDSE_HIP(hipStreamSynchronize(src->stream));
    TDSE_HIP(hipMemcpyAsync(dst->psi, src->psi, dst->total * dst->csize(), hipMemcpyDeviceToDevice,
                            dst->stream));
    return TDSE_OK;
}

int tdse_state_copy_kernel(tdse_plan* dst, const tdse_plan* src) {
    if (!dst || !src) return fail(TDSE_E_INVALID, "null plan");
    if (dst->total != src->total || dst->precision != src->precision)
        return fail(TDSE_E_INVALID, "shape or precision mismatch");
    int rc = set_device(dst->ctx);
    if (rc) return rc;
    TDSE_HIP(hipStreamSynchronize(src->stream));
    TDSE_HIP(tdse::launch_copy(dst->psi, src->psi, dst->total * dst->csize(), dst->stream));
    return TDSE_OK;
}

int tdse_state_device_ptr(tdse_plan* p, void** out) {
    int rc = check_plan(p);
    if (rc) return rc;
    if (!out) return fail(TDSE_E_INVALID, "null out");
    *out = p->psi;
    return TDSE_OK;
}

int tdse_step(tdse_plan* p, int nsteps) {
    int rc = check_plan(p);
    if (rc) return rc;
    if (nsteps < 0) return fail(TDSE_E_INVALID, "nsteps < 0");
    if (nsteps == 0) return TDSE_OK;
    if (!p->has_potential) return fail(TDSE_E_STATE, "tdse_plan_set_potential not called");
    if ((rc = set_device(p->ctx))) return rc;
    if (p->flags & TDSE_FLAG_UNFUSED_BASELINE) return run_baseline(p, nsteps);
    if (p->kind == kKindLattice) {
        if ((rc = run_lattice(p, nsteps))) return rc;
        p->time += nsteps * p->dt;
        return TDSE_OK;
    }
    if (p->bfield != 0.0) {
        if ((rc = run_magnetic(p, nsteps))) return rc;
        p->time += nsteps * p->dt;
        return TDSE_OK;
    }
    if ((rc = run_launches(p, schedule(p, nsteps)))) return rc;
    p->time += nsteps * p->dt;
    if (p->imag() && p->kind != kKindSingle) {
        // the last step's renormalisation, deferred by its measuring sweep
        TDSE_HIP(tdse::launch_scale_dev(p->fp64(), p->psi, p->total, &p->relax()->scale[nsteps & 1],
                                        p->stream));
    }
    return TDSE_OK;
}

int tdse_plan_set_drive(tdse_plan* p, const double axis[3], double amplitude, double omega) {
    int rc = check_plan(p);
    if (rc) return rc;
    if (!axis) return fail(TDSE_E_INVALID, "null axis");
    if (!std::isfinite(amplitude) || !std::isfinite(omega))
        return fail(TDSE_E_INVALID, "drive amplitude and frequency must be finite");
    if (p->imag() || (p->flags & TDSE_FLAG_UNFUSED_BASELINE) || p->bfield != 0.0 ||
        p->kind == kKindLattice)
        return fail(TDSE_E_UNSUPPORTED, "the drive applies to fused real-time plans without a field");
    for (int a = 0; a < 3; ++a)
        if (!std::isfinite(axis[a])) return fail(TDSE_E_INVALID, "drive axis must be finite");
    // The kick sweeps form the drive's angle (at most |E0| dt |axis . r| per
    // kick) with a short Cody-Waite reduction, exact and accurate for
    // arguments well below 2^20 pi / 2: drives whose angles reach 1e5 rad
    // anywhere in the box are refused, as for the lattice's link angles.
    double max_u = 0.0;
    for (int a = 0; a < 3; ++a) max_u += std::fabs(axis[a]) * std::max(std::fabs(p->lo[a]), std::fabs(p->hi[a]));
    if (std::fabs(amplitude) * std::fabs(p->dt) * max_u >= 1.0e5)
        return fail(TDSE_E_UNSUPPORTED, "drive angles |E0| dt |axis . r| must stay below 1e5 rad");
    for (int a = 0; a < 3; ++a) p->drive_axis[a] = axis[a];
    p->drive_amp = amplitude;
    p->drive_omega = omega;
    p->drive = amplitude != 0.0;
    return TDSE_OK;
}

int tdse_plan_set_magnetic(tdse_plan* p, double bfield, int axis) {
    int rc = check_plan(p);
    if (rc) return rc;
    if (!std::isfinite(bfield) || axis < 0 || axis > 2)
        return fail(TDSE_E_INVALID, "field must be finite, axis 0, 1 or 2");
    if (p->has_potential) return fail(TDSE_E_STATE, "set the field before the potential");
    if (bfield != 0.0) {
        if (p->imag() || (p->flags & TDSE_FLAG_UNFUSED_BASELINE) || p->drive ||
            p->kind == kKindLattice)
            return fail(TDSE_E_UNSUPPORTED, "the field applies to plain real-time plans");
        if (p->kind != kKind3D && !(p->kind == kKind2D && axis == 2 && p->n[2] == 1))
            return fail(TDSE_E_UNSUPPORTED, "the field needs a 3D grid (or a 2D one rotated about z)");
    }
    p->bfield = bfield;
    p->b_axis = axis;
    return TDSE_OK;
}

int tdse_lattice_set_uniform_field(tdse_plan* p, double b) {
    int rc = check_plan(p);
    if (rc) return rc;
    if (p->kind != kKindLattice) return fail(TDSE_E_INVALID, "not a lattice plan");
    if (!std::isfinite(b)) return fail(TDSE_E_INVALID, "field must be finite");
    // x links of row j: exp(-i B hx y_j) (core/src/lattice2d.ixx:96-105)
    const double a = -b * ((p->hi[0] - p->lo[0]) / p->n[0]);
    const double ymax = std::max(std::fabs(p->lo[1]), std::fabs(p->hi[1]));
    // the kernels reduce link angles by a short (Cody-Waite) reduction
    if (std::fabs(a) * ymax >= 1.0e5)
        return fail(TDSE_E_UNSUPPORTED, "link angles B hx y must stay below 1e5 rad");
    p->lat_mode = 1;
    p->lat_a = a;
    return TDSE_OK;
}

int tdse_lattice_set_solenoid(tdse_plan* p, double phi, double xs, double ys, int cut_up) {
    int rc = check_plan(p);
    if (rc) return rc;
    if (p->kind != kKindLattice) return fail(TDSE_E_INVALID, "not a lattice plan");
    if (!std::isfinite(phi) || !std::isfinite(xs) || !std::isfinite(ys))
        return fail(TDSE_E_INVALID, "solenoid arguments must be finite");
    // core/src/lattice2d.ixx:63-91: the cell holding (xs, ys); the string cut
    // runs up (rows above it, exp(-i phi)) or down (rows up to it, exp(+i phi))
    auto coord = [&](int a, int i) { return p->lo[a] + i * ((p->hi[a] - p->lo[a]) / p->n[a]); };
    int is = -1, js = -1;
    for (int i = 0; i + 1 < p->n[0]; ++i)
        if (coord(0, i) <= xs && xs < coord(0, i + 1)) is = i;
    for (int j = 0; j + 1 < p->n[1]; ++j)
        if (coord(1, j) <= ys && ys < coord(1, j + 1)) js = j;
    p->lat_mode = 0;  // links reset to 1 (a solenoid outside the lattice leaves them so)
    if (is < 0 || js < 0) return TDSE_OK;
    p->lat_mode = 2;
    p->lat_a = phi;
    p->lat_col = is;
    if (cut_up) {
        p->lat_j0 = js + 1;
        p->lat_j1 = p->n[1];
        p->lat_sign = -1.0;
    } else {
        p->lat_j0 = 0;
        p->lat_j1 = js + 1;
        p->lat_sign = 1.0;
    }
    return TDSE_OK;
}

int tdse_plan_set_time(tdse_plan* p, double t) {
    int rc = check_plan(p);
    if (rc) return rc;
    if (!std::isfinite(t)) return fail(TDSE_E_INVALID, "time must be finite");
    p->time = t;
    return TDSE_OK;
}

int tdse_plan_get_time(const tdse_plan* p, double* out) {
    if (!p || !out) return fail(TDSE_E_INVALID, "null argument");
    *out = p->time;
    return TDSE_OK;
}

int tdse_relax_deflated(tdse_plan* p, const tdse_plan* const* lower, int nlower, int nsteps) {
    int rc = check_plan(p);
    if (rc) return rc;
    if (!p->imag()) return fail(TDSE_E_INVALID, "not an imaginary-time plan");
    if (p->kind == kKindLattice) return fail(TDSE_E_UNSUPPORTED, "no deflation on the lattice");
    if (nsteps < 0 || nlower < 0 || (nlower > 0 && !lower))
        return fail(TDSE_E_INVALID, "bad deflation arguments");
    if (!p->has_potential) return fail(TDSE_E_STATE, "tdse_plan_set_potential not called");
    for (int j = 0; j < nlower; ++j) {
        const tdse_plan* q = lower[j];
        if (!q || q == p) return fail(TDSE_E_INVALID, "null or self lower state");
        if (q->ctx->device != p->ctx->device || q->precision != p->precision)
            return fail(TDSE_E_INVALID, "lower state on another device or in another precision");
        for (int a = 0; a < 3; ++a)
            if (q->n[a] != p->n[a]) return fail(TDSE_E_INVALID, "lower state of another shape");
    }
    if ((rc = set_device(p->ctx))) return rc;
    for (int j = 0; j < nlower; ++j) TDSE_HIP(hipStreamSynchronize(lower[j]->stream));
    const double dv = cell_volume(p);
    const tdse::ReduceArgs r = reduce_args(p);
    const std::vector<Launch> one = schedule(p, 1);
    // core/src/imaginary_time.ixx:112-134: relax step, Gram-Schmidt against
    // each lower state in order, normalise
    for (int s = 0; s < nsteps; ++s) {
        if ((rc = run_launches(p, one))) return rc;
        for (int j = 0; j < nlower; ++j) {
            TDSE_HIP(tdse::launch_inner(p->fp64(), lower[j]->psi, p->psi, p->total, p->partials,
                                        p->red_out, p->stream));
            TDSE_HIP(tdse::launch_subtract_projection(p->fp64(), p->psi, lower[j]->psi, p->total,
                                                      p->red_out, dv, p->stream));
        }
        TDSE_HIP(tdse::launch_reduce(p->fp64(), false, r, p->partials, p->red_out, p->stream));
        TDSE_HIP(tdse::launch_normalize_dev(p->fp64(), p->psi, p->total, p->red_out, dv,
                                            &p->relax()->norm_sq, p->stream));
    }
    return TDSE_OK;
}

int tdse_plan_relax_stats(tdse_plan* p, tdse_relax_stats* out) {
    int rc = check_plan(p);
    if (rc) return rc;
    if (!out) return fail(TDSE_E_INVALID, "null out");
    if (!p->imag()) return fail(TDSE_E_INVALID, "not an imaginary-time plan");
    if ((rc = set_device(p->ctx))) return rc;
    tdse::RelaxSlots h{};
    if ((rc = d2h(p, &h, p->relax(), sizeof(h)))) return rc;
    out->norm_sq = h.norm_sq;
    // the norm before renormalisation decays as exp(-2 E dtau)
    // (solver/src/vk_engine.ixx:3017-3023)
    out->energy = (h.norm_sq > 0.0 && p->dt > 0.0) ? -std::log(h.norm_sq) / (2.0 * p->dt) : 0.0;
    out->peak = h.norm_sq > 0.0 ? h.peak / h.norm_sq : 0.0;
    return TDSE_OK;
}

int tdse_plan_sync(tdse_plan* p, int timeout_ms) {
    int rc = check_plan(p);
    if (rc) return rc;
    if ((rc = set_device(p->ctx))) return rc;
    if (timeout_ms < 0) {
        TDSE_HIP(hipStreamSynchronize(p->stream));
        return TDSE_OK;
    }
    const auto t0 = std::chrono::steady_clock::now();
    for (;;) {
        hipError_t e = hipStreamQuery(p->stream);
        if (e == hipSuccess) return TDSE_OK;
        if (e != hipErrorNotReady) return fail(TDSE_E_DEVICE, hipGetErrorString(e));
        const auto ms = std::chrono::duration_cast<std::chrono::milliseconds>(
                            std::chrono::steady_clock::now() - t0)
                            .count();
        if (ms >= timeout_ms) return fail(TDSE_E_TIMEOUT, "stream wait timed out");
        std::this_thread::yield();
    }
}

int tdse_plan_stream(tdse_plan* p, void** out) {
    int rc = check_plan(p);
    if (rc) return rc;
    if (!out) return fail(TDSE_E_INVALID, "null out");
    *out = p->stream;
    return TDSE_OK;
}

namespace {

// Var(H) = ||(H - <H>) psi||^2 / ||psi||^2 of a 1D plan's state
// (core/src/observables.ixx:92-119): T psi = IFFT(0.5 k^2 / m FFT(psi)) in the
// plan's scratch array with the step path's sweep kernels (a line <= 1024:
// one FFT-K-IFFT sweep; a four-step line: column FFT and twiddle, row
// FFT-K-IFFT and conjugate twiddle, column IFFT), then H psi = T psi + V psi
// and the two fixed-order reductions of the residual form.
int energy_variance(tdse_plan* p, tdse_observables* out) {
    if (p->kind != kKindSingle && p->kind != kKindFourStep)
        return fail(TDSE_E_UNSUPPORTED, "the energy variance covers 1D grids (Field1D)");
    int rc;
    if (!p->scratch && (rc = alloc(p, &p->scratch, p->total * p->csize()))) return rc;
    if (!p->kvar) {
        if ((rc = alloc(p, &p->kvar, p->total * 16))) return rc;
        const Axis& ax = p->sq[0];
        const double dk = dk_of(ax);
        std::vector<double> t(2 * static_cast<size_t>(ax.n));
        for (int j = 0; j < ax.n; ++j) {
            const double k = dk * ((2 * j < ax.n) ? j : j - ax.n);
            t[2 * j] = 0.5 * k * k / p->mass;
            t[2 * j + 1] = 0.0;
        }
        if ((rc = h2d(p, p->kvar, t.data(), t.size() * 8))) {
            release(p, p->kvar, p->total * 16);
            return rc;
        }
    }
    TDSE_HIP(hipMemcpyAsync(p->scratch, p->psi, p->total * p->csize(), hipMemcpyDeviceToDevice,
                            p->stream));
    Geometry g = geometry(p, p->scratch);
    std::vector<Launch> L;
    if (p->kind == kKindSingle) {
        SweepArgs a = g.single;
        a.kt[0] = a.kt[1] = a.kt[2] = nullptr;
        a.kt[2] = p->kvar;
        a.n_rows = 1;
        a.n_total = p->sq[0].n;
        L.push_back({tdse::C_KIN_TW, p->sq[0].n, a, launcher(p, p->sq[0].n, tdse::C_KIN_TW), 0});
    } else {
        SweepArgs row = g.row;
        row.kt[2] = p->kvar;
        L.push_back({tdse::S_FWD_TW, p->n2, g.col, launcher(p, p->n2, tdse::S_FWD_TW), 0});
        L.push_back({tdse::C_KIN_TW, p->n1, row, launcher(p, p->n1, tdse::C_KIN_TW), 0});
        L.push_back({tdse::S_INV, p->n2, g.col, launcher(p, p->n2, tdse::S_INV), 0});
    }
    if ((rc = run_launches(p, L))) return rc;
    double* sums = p->red_out + 40;  // 2 kReduceQ doubles
    static_assert(40 + 2 * tdse::kReduceQ <= 64, "variance sums fit the reduction output");
    TDSE_HIP(tdse::launch_energy_variance(p->fp64(), p->psi, p->scratch,
                                          p->has_potential ? p->v : nullptr, p->total, p->partials,
                                          sums, p->stream));
    double h[2 * tdse::kReduceQ];
    if ((rc = d2h(p, h, sums, sizeof(h)))) return rc;
    out->energy_variance = h[0] > 0.0 ? h[tdse::kReduceQ] / h[0] : 0.0;
    return TDSE_OK;
}

}  // namespace

int tdse_observe(tdse_plan* p, unsigned what, tdse_observables* out) {
    int rc = check_plan(p);
    if (rc) return rc;
    if (!out) return fail(TDSE_E_INVALID, "null out");
    if ((what & TDSE_OBS_VARIANCE) && p->kind != kKindSingle && p->kind != kKindFourStep)
        return fail(TDSE_E_UNSUPPORTED, "the energy variance covers 1D grids (Field1D)");
    if ((rc = set_device(p->ctx))) return rc;
    std::memset(out, 0, sizeof(*out));
    tdse::ReduceArgs r = reduce_args(p);
    TDSE_HIP(tdse::launch_reduce(p->fp64(), false, r, p->partials, p->red_out, p->stream));
    auto ratio = [](double num, double den) { return den > 0.0 ? num / den : 0.0; };
    const bool want_k = (what & (TDSE_OBS_ENERGY | TDSE_OBS_MOMENTUM)) != 0;
    // original axis of each squeezed axis
    int orig[3] = {0, 0, 0};
    for (int a = 0, q = 0; a < 3; ++a)
        if (p->n[a] > 1) orig[q++] = a;
    // Grids swept by lines (not the lattice, not a four-step 1D grid): the
    // k-space sweeps are queued behind the position-space reduction (its final
    // pass has consumed the partials they reuse) and both results come back
    // in one copy.  Separable along axes: sum_k k_a^2 |psi_k|^2 / sum_k
    // |psi_k|^2 equals the same ratio over the 1D spectra of the lines along
    // axis a (the transforms along the other axes are unitary up to a
    // constant), so one read-only line-FFT sweep per axis gives <k_a> and
    // <k_a^2>.
    const bool line_k = want_k && p->kind != kKindLattice && p->kind != kKindFourStep;
    const int axes = static_cast<int>(p->sq.size());
    constexpr int kObsAt = 24;  // offset of the per-axis spectra sums in red_out
    static_assert(tdse::kReduceQ <= kObsAt, "position sums and spectra sums overlap");
    if (line_k) {
        const Geometry g = geometry(p, p->psi);
        for (int a = 0; a < axes; ++a) {
            SweepArgs args;
            int mode, n = p->sq[a].n;
            if (p->kind == kKindSingle) {
                args = g.single;
                mode = tdse::C_OBS;
            } else if (a == 0) {
                args = g.x;
                mode = tdse::C_OBS;
            } else if (a == 1) {
                args = g.yf;
                mode = tdse::S_OBS;
            } else {
                args = g.z;
                mode = tdse::S_OBS;
            }
            args.dk = dk_of(p->sq[a]);
            args.partials = p->partials;
            const tdse::LaunchFn fn = launcher(p, n, mode);
            if (!fn) return fail(TDSE_E_UNSUPPORTED, "no observing kernel for this length");
            TDSE_HIP(fn(args, tdse::kObsMaxBlocks, p->stream));
            const long long tiles = tiles_of(p->precision, n, mode, args);
            const int blocks = static_cast<int>(std::min<long long>(tiles, tdse::kObsMaxBlocks));
            TDSE_HIP(tdse::launch_reduce_final3(p->partials, blocks, p->red_out + kObsAt + 3 * a,
                                                p->stream));
        }
    }
    double host[kObsAt + 9] = {};
    const size_t count = line_k ? static_cast<size_t>(kObsAt + 3 * axes) : tdse::kReduceQ;
    TDSE_HIP(hipMemcpyAsync(host, p->red_out, count * sizeof(double), hipMemcpyDeviceToHost,
                            p->stream));
    TDSE_HIP(hipStreamSynchronize(p->stream));
    const double cell = r.h[0] * r.h[1] * r.h[2];
    out->sum_abs2 = host[0];
    out->norm = host[0] * cell;
    out->position[0] = ratio(host[1], host[0]);
    out->position[1] = ratio(host[2], host[0]);
    out->position[2] = ratio(host[3], host[0]);
    out->potential = ratio(host[4], host[0]);
    for (int a = 0; a < 3; ++a) {
        const double m = ratio(host[1 + a], host[0]);
        out->sigma[a] = std::sqrt(std::max(0.0, ratio(host[6 + a], host[0]) - m * m));
    }
    if (what & TDSE_OBS_PEAK) {
        // the sums above are on the host: the peak reuses their scratch
        TDSE_HIP(tdse::launch_peak(p->fp64(), p->psi, p->total, p->partials, p->red_out, p->stream));
        if ((rc = d2h(p, &out->peak, p->red_out, sizeof(double)))) return rc;
    }
    if (!want_k) return (what & TDSE_OBS_VARIANCE) ? energy_variance(p, out) : TDSE_OK;
    if (p->kind == kKindLattice) {
        // <H> of the lattice Hamiltonian (hops + onsite, core/src/lattice2d.ixx:251-274)
        if (what & TDSE_OBS_MOMENTUM)
            return fail(TDSE_E_UNSUPPORTED, "no momentum observable on the lattice");
        double* dst = p->red_out + kObsAt;
        TDSE_HIP(tdse::launch_lattice_energy(p->fp64(), lattice_args(p), p->partials, dst, p->stream));
        double hk[3];
        if ((rc = d2h(p, hk, dst, sizeof(hk)))) return rc;
        out->energy = ratio(hk[0], hk[1]);
        out->kinetic = out->energy - out->potential;
        return TDSE_OK;
    }
    if (p->kind == kKindFourStep) {
        // the 1D transform of a four-step grid, in a scratch copy
        if (!p->scratch && (rc = alloc(p, &p->scratch, p->total * p->csize()))) return rc;
        TDSE_HIP(hipMemcpyAsync(p->scratch, p->psi, p->total * p->csize(), hipMemcpyDeviceToDevice,
                                p->stream));
        Geometry g = geometry(p, p->scratch);
        std::vector<Launch> L;
        L.push_back({tdse::S_FWD_TW, p->n2, g.col, launcher(p, p->n2, tdse::S_FWD_TW), 0});
        L.push_back({tdse::C_FWD, p->n1, g.row, launcher(p, p->n1, tdse::C_FWD), 0});
        if ((rc = run_launches(p, L))) return rc;
        tdse::ReduceArgs rk = r;
        rk.psi = p->scratch;
        rk.v = nullptr;
        rk.sq[0] = p->n1;
        rk.sq[1] = p->n2;
        rk.four_step_rows = p->n2;
        double hk[tdse::kReduceQ];
        TDSE_HIP(tdse::launch_reduce(p->fp64(), true, rk, p->partials, p->red_out, p->stream));
        TDSE_HIP(hipMemcpyAsync(hk, p->red_out, sizeof(hk), hipMemcpyDeviceToHost, p->stream));
        TDSE_HIP(hipStreamSynchronize(p->stream));
        out->kinetic = ratio(hk[1], hk[0]);
        out->momentum[orig[0]] = ratio(hk[2], hk[0]);
    } else {
        const double* hk = host + kObsAt;
        double t = 0.0;
        for (int a = 0; a < axes; ++a) {
            t += 0.5 * ratio(hk[3 * a + 2], hk[3 * a]) / p->mass;
            out->momentum[orig[a]] = ratio(hk[3 * a + 1], hk[3 * a]);
        }
        out->kinetic = t;
    }
    out->energy = out->potential + out->kinetic;
    if (what & TDSE_OBS_VARIANCE) return energy_variance(p, out);
    return TDSE_OK;
}

namespace {

// |psi|^2 summed over the box r.box_lo <= coordinate < r.box_hi, times the
// cell volume (slot 5 of the x-space reduction).
int region_probability(tdse_plan* p, tdse::ReduceArgs& r, double* out) {
    int rc = set_device(p->ctx);
    if (rc) return rc;
    r.psi = p->psi;
    r.v = nullptr;
    r.total = p->total;
    for (int a = 0; a < 3; ++a) {
        r.n[a] = p->n[a];
        r.lo[a] = p->lo[a];
        r.h[a] = (p->hi[a] - p->lo[a]) / p->n[a];
    }
    r.mass = p->mass;
    double host[tdse::kReduceQ];
    TDSE_HIP(tdse::launch_reduce(p->fp64(), false, r, p->partials, p->red_out, p->stream));
    TDSE_HIP(hipMemcpyAsync(host, p->red_out, sizeof(host), hipMemcpyDeviceToHost, p->stream));
    TDSE_HIP(hipStreamSynchronize(p->stream));
    *out = host[5] * (r.h[0] * r.h[1] * r.h[2]);
    return TDSE_OK;
}

}  // namespace

int tdse_probability_in_slab(tdse_plan* p, int axis, double lo, double hi, double* out) {
    int rc = check_plan(p);
    if (rc) return rc;
    if (!out || axis < 0 || axis > 2) return fail(TDSE_E_INVALID, "bad axis or null out");
    tdse::ReduceArgs r{};
    tdse::open_box(r);
    r.box_lo[axis] = lo;
    r.box_hi[axis] = hi;
    return region_probability(p, r, out);
}

int tdse_probability_in_box(tdse_plan* p, const double lo[3], const double hi[3], double* out) {
    int rc = check_plan(p);
    if (rc) return rc;
    if (!out || !lo || !hi) return fail(TDSE_E_INVALID, "null argument");
    tdse::ReduceArgs r{};
    tdse::open_box(r);
    for (int a = 0; a < 3; ++a) {
        if (std::isnan(lo[a]) || std::isnan(hi[a])) return fail(TDSE_E_INVALID, "NaN box bound");
        if (p->n[a] > 1) {  // axes of extent 1 are not restricted
            r.box_lo[a] = lo[a];
            r.box_hi[a] = hi[a];
        }
    }
    return region_probability(p, r, out);
}

int tdse_state_scale(tdse_plan* p, double s) {
    int rc = check_plan(p);
    if (rc) return rc;
    if ((rc = set_device(p->ctx))) return rc;
    TDSE_HIP(tdse::launch_scale(p->fp64(), p->psi, p->total, s, p->stream));
    return TDSE_OK;
}

int tdse_profile_step(tdse_plan* p, int nsteps, int max_slots, double* ms_per_sweep, int* launches,
                      int* n_slots) {
    int rc = check_plan(p);
    if (rc) return rc;
    if (p->kind == kKindLattice) return fail(TDSE_E_UNSUPPORTED, "the lattice has no sweep slots");
    if (nsteps < 1 || !ms_per_sweep || !launches || !n_slots)
        return fail(TDSE_E_INVALID, "bad profiling arguments");
    if (p->flags & TDSE_FLAG_UNFUSED_BASELINE)
        return fail(TDSE_E_UNSUPPORTED, "per-sweep profiling covers the fused schedule");
    if (!p->has_potential) return fail(TDSE_E_STATE, "tdse_plan_set_potential not called");
    if ((rc = set_device(p->ctx))) return rc;
    const std::vector<Launch> L = schedule(p, nsteps);
    std::vector<hipEvent_t> ev(2 * L.size());
    for (auto& e : ev) TDSE_HIP(hipEventCreate(&e));
    int status = TDSE_OK;
    // Sweeps record their events at their kernel's own start and end
    // (tdse::g_launch_timing); the norm finish of a relax step, launched
    // elsewhere, between events recorded around it.
    for (size_t i = 0; i < L.size() && status == TDSE_OK; ++i) {
        if (!L[i].fn) {
            status = fail(TDSE_E_UNSUPPORTED, "no kernel");
            break;
        }
        hipError_t e = hipSuccess;
        if (L[i].mode == kFinishMode) {
            if ((e = hipEventRecord(ev[2 * i], p->stream)) == hipSuccess &&
                (e = L[i].fn(L[i].args, p->ctx->max_blocks, p->stream)) == hipSuccess)
                e = hipEventRecord(ev[2 * i + 1], p->stream);
        } else {
            tdse::g_launch_timing = tdse::LaunchTiming{ev[2 * i], ev[2 * i + 1]};
            e = L[i].fn(L[i].args, p->ctx->max_blocks, p->stream);
            tdse::g_launch_timing = tdse::LaunchTiming{};
        }
        if (e != hipSuccess) status = fail(TDSE_E_DEVICE, "launch failed while profiling");
    }
    if (status == TDSE_OK && hipStreamSynchronize(p->stream) != hipSuccess)
        status = fail(TDSE_E_DEVICE, "sync failed while profiling");
    const int slots = slot_count(p);
    *n_slots = slots;
    for (int s = 0; s < std::min(slots, max_slots); ++s) {
        ms_per_sweep[s] = 0.0;
        launches[s] = 0;
    }
    if (status == TDSE_OK) {
        for (size_t i = 0; i < L.size(); ++i) {
            float ms = 0.f;
            if (hipEventElapsedTime(&ms, ev[2 * i], ev[2 * i + 1]) != hipSuccess) continue;
            if (L[i].slot >= 0 && L[i].slot < max_slots) {
                ms_per_sweep[L[i].slot] += ms;
                if (L[i].mode != kFinishMode) launches[L[i].slot] += 1;
            }
        }
    }
    for (auto& e : ev) (void)hipEventDestroy(e);
    return status;
}

const char* tdse_plan_sweep_name(const tdse_plan* p, int slot) {
    if (!p || slot < 0 || slot >= slot_count(p)) return "";
    switch (p->kind) {
        case kKind3D: return kSlots3D[slot];
        case kKind2D: return kSlots2D[slot];
        case kKindSingle: return kSlotsSingle[slot];
        default: return kSlotsFour[slot];
    }
}

int tdse_plan_get_variants(const tdse_plan* p, int max_slots, int* variants, int* n_slots) {
    if (!p || !variants || !n_slots || max_slots < 0) return fail(TDSE_E_INVALID, "null argument");
    const int slots = slot_count(p);
    *n_slots = slots;
    for (int s = 0; s < std::min(slots, max_slots); ++s) variants[s] = p->variant[s];
    return TDSE_OK;
}

int tdse_plan_set_variants(tdse_plan* p, const int* variants, int n_slots) {
    if (!p || !variants) return fail(TDSE_E_INVALID, "null argument");
    if (p->kind == kKindLattice) return fail(TDSE_E_UNSUPPORTED, "the lattice has no sweep slots");
    const int slots = slot_count(p);
    if (n_slots != slots) return fail(TDSE_E_INVALID, "expected one variant per schedule slot");
    const Geometry g = geometry(p, p->psi);
    for (int s = 0; s < slots; ++s) {
        const int v = variants[s];
        if (v < 0 || v >= tdse::kVariants) return fail(TDSE_E_INVALID, "variant out of range");
        const SlotSweep w = slot_sweep(p, g, s);
        if (!launcher(p, w.n, w.mode, v))
            return fail(TDSE_E_UNSUPPORTED, "variant " + std::to_string(v) + " not compiled for slot " +
                                                std::to_string(s));
    }
    for (int s = 0; s < slots; ++s) p->variant[s] = variants[s];
    return TDSE_OK;
}

namespace {

// Timing helpers of the autotuner.  Everything runs on the plan's scratch
// array, refreshed from the state before each measurement; the state is
// never written.
struct Tuner {
    tdse_plan* p;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    size_t bytes = 0;

    int refresh() {
        TDSE_HIP(hipMemcpyAsync(p->scratch, p->psi, bytes, hipMemcpyDeviceToDevice, p->stream));
        return TDSE_OK;
    }

    // Milliseconds per steady-state step of the current variants: a schedule
    // of `k` steps, timed from after its first sweep to before its last; the
    // best of `rounds` runs.
    int per_step(int k, double* out, int rounds = 3) {
        const std::vector<Launch> L = schedule(p, k, p->scratch);
        for (const Launch& l : L)
            if (!l.fn) return fail(TDSE_E_UNSUPPORTED, "no kernel for a tuned sweep");
        double best = 1e30;
        for (int round = 0; round < rounds; ++round) {
            int rc = refresh();
            if (rc) return rc;
            TDSE_HIP(L.front().fn(L.front().args, p->ctx->max_blocks, p->stream));
            TDSE_HIP(hipEventRecord(e0, p->stream));
            for (size_t i = 1; i + 1 < L.size(); ++i)
                TDSE_HIP(L[i].fn(L[i].args, p->ctx->max_blocks, p->stream));
            TDSE_HIP(hipEventRecord(e1, p->stream));
            TDSE_HIP(L.back().fn(L.back().args, p->ctx->max_blocks, p->stream));
            TDSE_HIP(hipStreamSynchronize(p->stream));
            float ms = 0.f;
            TDSE_HIP(hipEventElapsedTime(&ms, e0, e1));
            best = std::min(best, static_cast<double>(ms) / k);
        }
        *out = best;
        return TDSE_OK;
    }

    // Per-step times of variant sets `a` and `b`, measured in alternation
    // (a b a b ...) so that a drift of clocks or temperature during tuning
    // falls on both; the best of `pairs` runs each.
    int compare(const int* a, const int* b, int k, int pairs, double* ta, double* tb) {
        *ta = *tb = 1e30;
        for (int i = 0; i < pairs; ++i) {
            for (int side = 0; side < 2; ++side) {
                const int* set = side ? b : a;
                for (int s = 0; s < kMaxSlots; ++s) p->variant[s] = set[s];
                double t = 0.0;
                const int rc = per_step(k, &t, 1);
                if (rc) return rc;
                double* dst = side ? tb : ta;
                *dst = std::min(*dst, t);
            }
        }
        return TDSE_OK;
    }

    // Milliseconds per launch of one sweep repeated `k` times (after two
    // untimed launches); the best of three runs.
    int per_launch(tdse::LaunchFn fn, const SweepArgs& a, int k, double* out) {
        double best = 1e30;
        for (int round = 0; round < 3; ++round) {
            int rc = refresh();
            if (rc) return rc;
            for (int i = 0; i < 2; ++i) TDSE_HIP(fn(a, p->ctx->max_blocks, p->stream));
            TDSE_HIP(hipEventRecord(e0, p->stream));
            for (int i = 0; i < k; ++i) TDSE_HIP(fn(a, p->ctx->max_blocks, p->stream));
            TDSE_HIP(hipEventRecord(e1, p->stream));
            TDSE_HIP(hipStreamSynchronize(p->stream));
            float ms = 0.f;
            TDSE_HIP(hipEventElapsedTime(&ms, e0, e1));
            best = std::min(best, static_cast<double>(ms) / k);
        }
        *out = best;
        return TDSE_OK;
    }
};

// Slots swept once per step (tuned on whole steps) and once per call (tuned
// per launch).
void slot_roles(const tdse_plan* p, std::vector<int>* per_step, std::vector<int>* per_call) {
    switch (p->kind) {
        case kKind3D:
            *per_step = {1, 2, 3, 4};
            *per_call = {0, 5};
            break;
        case kKindSingle:
            *per_step = {};
            *per_call = {};
            break;
        default:
            *per_step = {1, 2};
            *per_call = {0, 3};
    }
}

int autotune_impl(tdse_plan* p, Tuner& t, int reps, int* chosen) {
    std::vector<int> per_step, per_call;
    slot_roles(p, &per_step, &per_call);
    const Geometry g = geometry(p, p->scratch);
    for (int s = 0; s < kMaxSlots; ++s) p->variant[s] = 0;
    int rc;
    if (!per_step.empty()) {
        // enough steps for about 2 ms of work per measurement
        int k = reps;
        double ms = 0.0;
        if ((rc = t.per_step(k, &ms))) return rc;
        if (ms * k < 2.0) k = static_cast<int>(std::min(256.0, std::ceil(2.0 / std::max(ms, 1e-4))));
        // one pass of coordinate descent over the per-step slots; each
        // candidate is timed in alternation with the current choice and
        // must win by 2%
        int cur[kMaxSlots] = {};
        for (int s : per_step) {
            const SlotSweep w = slot_sweep(p, g, s);
            for (int v = 1; v < tdse::kVariants; ++v) {
                if (!launcher(p, w.n, w.mode, v)) continue;
                int cand[kMaxSlots];
                std::copy(cur, cur + kMaxSlots, cand);
                cand[s] = v;
                double t_cur = 0.0, t_cand = 0.0;
                if ((rc = t.compare(cur, cand, k, 3, &t_cur, &t_cand))) return rc;
                if (t_cand < 0.98 * t_cur) std::copy(cand, cand + kMaxSlots, cur);
            }
        }
        // the set found must still beat the default kernels by 2% in a
        // longer alternation, else the default is kept
        const int zero[kMaxSlots] = {};
        if (!std::equal(cur, cur + kMaxSlots, zero)) {
            double t0 = 0.0, t1 = 0.0;
            if ((rc = t.compare(zero, cur, k, 5, &t0, &t1))) return rc;
            if (!(t1 < 0.98 * t0)) std::fill(cur, cur + kMaxSlots, 0);
        }
        for (int s = 0; s < kMaxSlots; ++s) p->variant[s] = cur[s];
    }
    for (int s : per_call) {
        const SlotSweep w = slot_sweep(p, g, s);
        double t0 = 0.0;
        if ((rc = t.per_launch(launcher(p, w.n, w.mode, 0), w.args, reps, &t0))) return rc;
        double best = t0;
        for (int v = 1; v < tdse::kVariants; ++v) {
            const tdse::LaunchFn fn = launcher(p, w.n, w.mode, v);
            if (!fn) continue;
            double t_v = 0.0;
            if ((rc = t.per_launch(fn, w.args, reps, &t_v))) return rc;
            if (t_v < 0.98 * best) {
                best = t_v;
                p->variant[s] = v;
            }
        }
    }
    for (int s = 0; s < kMaxSlots; ++s) chosen[s] = p->variant[s];
    return TDSE_OK;
}

}  // namespace

int tdse_plan_autotune(tdse_plan* p, int reps, int max_slots, int* variants, int* n_slots) {
    int rc = check_plan(p);
    if (rc) return rc;
    if (p->kind == kKindLattice) return fail(TDSE_E_UNSUPPORTED, "the lattice has no sweep slots");
    if (reps < 1 || reps > 64 || !variants || !n_slots || max_slots < 0)
        return fail(TDSE_E_INVALID, "bad autotune arguments");
    if (!p->has_potential) return fail(TDSE_E_STATE, "tdse_plan_set_potential not called");
    if (p->flags & TDSE_FLAG_UNFUSED_BASELINE)
        return fail(TDSE_E_UNSUPPORTED, "the unfused baseline has no kernel variants");
    if ((rc = set_device(p->ctx))) return rc;
    const int slots = slot_count(p);
    *n_slots = slots;
    if (!p->scratch && (rc = alloc(p, &p->scratch, p->total * p->csize()))) return rc;
    int saved[kMaxSlots], chosen[kMaxSlots];
    for (int s = 0; s < kMaxSlots; ++s) saved[s] = p->variant[s];
    // timing runs the measuring sweeps of an imaginary-time plan, which
    // write its statistics: keep them
    tdse::RelaxSlots kept{};
    if (p->imag() && (rc = d2h(p, &kept, p->relax(), sizeof(kept)))) return rc;
    Tuner t{p};
    t.bytes = static_cast<size_t>(p->total) * p->csize();
    if (hipEventCreate(&t.e0) != hipSuccess || hipEventCreate(&t.e1) != hipSuccess) {
        if (t.e0) (void)hipEventDestroy(t.e0);
        return fail(TDSE_E_DEVICE, "event creation failed");
    }
    rc = autotune_impl(p, t, reps, chosen);
    (void)hipEventDestroy(t.e0);
    (void)hipEventDestroy(t.e1);
    if (p->imag()) {
        const int rc2 = h2d(p, p->relax(), &kept, sizeof(kept));
        if (!rc) rc = rc2;
    }
    if (rc) {
        for (int s = 0; s < kMaxSlots; ++s) p->variant[s] = saved[s];
        return rc;
    }
    for (int s = 0; s < std::min(slots, max_slots); ++s) variants[s] = chosen[s];
    return TDSE_OK;
}

int tdse_plan_step_bytes(const tdse_plan* p, uint64_t* out) {
    if (!p || !out) return fail(TDSE_E_INVALID, "null argument");
    int sweeps = 1;
    if (p->kind == kKind3D) sweeps = 4;
    else if (p->kind == kKind2D || p->kind == kKindFourStep) sweeps = 2;
    const uint64_t N = static_cast<uint64_t>(p->total);
    if (p->kind == kKindLattice) {
        // step tiles, two steps per launch: a 48 x 44 window (state, onsite
        // phase, mask) read per 32 x 32 sites written; a call's opening row
        // sweep is not counted
        const uint64_t in = p->csize() + p->rsize() + (p->mask ? p->rsize() : 0);
        *out = N * (33 * in + 16 * p->csize()) / 32;
        return TDSE_OK;
    }
    if (p->bfield != 0.0) {
        // the sweeps of one unmerged Strang step (two kick sweeps read the
        // phase) and six shear sweeps, mask with the last shear.  Field along
        // z: the x shears next to the core step join its x sweeps, and
        // without a mask the closing and opening x shears of neighbouring
        // steps are one sweep.
        const int core = p->kind == kKind3D ? 5 : 3;
        int sweeps = core + 6;
        if (p->b_axis == 2) sweeps = core + (p->mask ? 4 : 3);
        *out = N * (sweeps * 2 * p->csize() + 2 * p->rsize() + (p->mask ? p->rsize() : 0));
        return TDSE_OK;
    }
    if (p->flags & TDSE_FLAG_UNFUSED_BASELINE) {
        // 3 table passes (48 B), 2 x (axes) FFT sweeps (32 B), 2 conj passes (32 B), mask (40 B)
        const int axes = p->kind == kKind3D ? 3 : 2;
        *out = N * (3 * 48 + 2 * axes * 32 + 2 * 32 + (p->mask ? 40 : 0));
        return TDSE_OK;
    }
    // the x sweeps of a plan with mask ramps read no mask array
    const bool mask_array = p->mask && !(p->mask_x && mask_axes_apply(p));
    *out = N * (sweeps * 2 * p->csize() + p->rsize() + (mask_array ? p->rsize() : 0));
    return TDSE_OK;
}

int tdse_batch_create(tdse_ctx* ctx, tdse_plan* const* plans, int count, tdse_batch** out) {
    if (!ctx || !plans || count < 1 || !out) return fail(TDSE_E_INVALID, "bad batch arguments");
    *out = nullptr;
    int rc = set_device(ctx);
    if (rc) return rc;
    for (int i = 0; i < count; ++i) {
        if (!plans[i] || plans[i]->ctx != ctx) return fail(TDSE_E_INVALID, "plan of another context");
        if (!plans[i]->has_potential) return fail(TDSE_E_STATE, "plan without potential");
        if (plans[i]->flags & TDSE_FLAG_UNFUSED_BASELINE)
            return fail(TDSE_E_INVALID, "the unfused baseline does not batch");
        if (plans[i]->imag())
            return fail(TDSE_E_UNSUPPORTED, "imaginary-time plans do not batch");
        if (plans[i]->drive || plans[i]->bfield != 0.0 || plans[i]->kind == kKindLattice)
            return fail(TDSE_E_UNSUPPORTED, "driven, magnetic and lattice plans do not batch");
        for (int j = 0; j < i; ++j)
            if (plans[j] == plans[i]) return fail(TDSE_E_INVALID, "plan listed twice");
    }
    tdse_batch* b = new tdse_batch();
    b->ctx = ctx;
    b->plans.assign(plans, plans + count);
    auto bail = [&](int code) {
        tdse_batch_destroy(b);
        return code;
    };
    if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess) {
        (void)hipGetLastError();
        b->stream = nullptr;
        return bail(fail(TDSE_E_DEVICE, "stream creation failed"));
    }
    register_stream(ctx, b->stream);
    if (hipEventCreateWithFlags(&b->fork, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&b->done, hipEventDisableTiming) != hipSuccess) {
        (void)hipGetLastError();
        return bail(fail(TDSE_E_DEVICE, "event creation failed"));
    }
    for (int s = 0; s < tdse_batch::kSide; ++s) {
        if (hipStreamCreateWithFlags(&b->side[s], hipStreamNonBlocking) != hipSuccess ||
            hipEventCreateWithFlags(&b->join[s], hipEventDisableTiming) != hipSuccess) {
            (void)hipGetLastError();
            return bail(fail(TDSE_E_DEVICE, "stream/event creation failed"));
        }
    }
    // Units: members of one class (precision, geometry, line lengths) join a
    // unit while its one-step working set stays within a quarter of the
    // Infinity Cache (one lane's share), so a unit's state stays cached
    // between its sweeps; a call runs each unit's steps back to back on one
    // lane.  Units go to the lanes in the order of their first member.
    for (int i = 0; i < count; ++i) {
        tdse_plan* p = b->plans[i];
        const uint64_t ws = step_working_set(p);
        BatchUnit* u = nullptr;
        for (BatchUnit& cand : b->units)
            if (same_class(b->plans[cand.members.front()], p) &&
                cand.bytes + ws <= static_cast<uint64_t>(TDSE_BATCH_UNIT_BYTES)) {
                u = &cand;
                break;
            }
        if (!u) {
            b->units.emplace_back();
            u = &b->units.back();
        }
        u->members.push_back(i);
        u->bytes += ws;
    }
    for (size_t k = 0; k < b->units.size(); ++k) {
        BatchUnit& u = b->units[k];
        u.lane = static_cast<int>(k % (tdse_batch::kSide + 1));
        std::vector<BatchRow> init_rows, mid_rows[4], last_rows[4];
        for (int i : u.members) {
            tdse_plan* p = b->plans[i];
            if (p->kind == kKindSingle) {
                // every step of a call in one block: nsteps is set per call
                Geometry g = geometry(p, p->psi);
                init_rows.push_back({p, p->sq[0].n, tdse::C_SINGLE, 0, g.single});
                continue;
            }
            int n, mode, var;
            SweepArgs a;
            if (plan_init(p, &n, &mode, &a, &var)) init_rows.push_back({p, n, mode, var, a});
            for (const RoundSweep& r : plan_round(p)) {
                mid_rows[r.sub].push_back({p, r.n, r.mode_mid, r.var_mid, r.args});
                last_rows[r.sub].push_back({p, r.n, r.mode_last, r.var_last, r.args});
            }
        }
        if ((rc = build_groups(b, u.init, init_rows))) return bail(rc);
        for (int sub = 0; sub < 4; ++sub) {
            if ((rc = build_groups(b, u.mid[sub], mid_rows[sub]))) return bail(rc);
            if ((rc = build_groups(b, u.last[sub], last_rows[sub]))) return bail(rc);
        }
    }
    *out = b;
    return TDSE_OK;
}

int tdse_batch_destroy(tdse_batch* b) {
    if (!b) return fail(TDSE_E_INVALID, "null batch");
    (void)hipSetDevice(b->ctx->device);
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    auto free_groups = [](std::vector<BatchGroup>& gs) {
        for (auto& g : gs) {
            if (g.d_items) (void)hipFree(g.d_items);
            if (g.d_first) (void)hipFree(g.d_first);
        }
        gs.clear();
    };
    for (BatchUnit& u : b->units) {
        free_groups(u.init);
        for (int sub = 0; sub < 4; ++sub) {
            free_groups(u.mid[sub]);
            free_groups(u.last[sub]);
        }
    }
    for (int s = 0; s < tdse_batch::kSide; ++s) {
        if (b->side[s]) (void)hipStreamDestroy(b->side[s]);
        if (b->join[s]) (void)hipEventDestroy(b->join[s]);
    }
    if (b->fork) (void)hipEventDestroy(b->fork);
    if (b->done) (void)hipEventDestroy(b->done);
    if (b->stream) {
        unregister_stream(b->ctx, b->stream);
        (void)hipStreamDestroy(b->stream);
    }
    delete b;
    return TDSE_OK;
}

int tdse_batch_step(tdse_batch* b, int nsteps, int use_graph) {
    if (!b) return fail(TDSE_E_INVALID, "null batch");
    // Replaying stream-captured hipGraphs of the batch rounds crashed the host
    // process inside tdse_batch_step with no cause shown in this library; the
    // path is switched off until one is.
    if (use_graph) return fail(TDSE_E_UNSUPPORTED, "captured graph replay is switched off");
    if (nsteps < 0) return fail(TDSE_E_INVALID, "nsteps < 0");
    if (nsteps == 0) return TDSE_OK;
    int rc = set_device(b->ctx);
    if (rc) return rc;
    // the plans' own queues must be drained before the batch touches their state
    for (tdse_plan* p : b->plans) TDSE_HIP(hipStreamSynchronize(p->stream));
    // fork: every lane starts after the main stream's earlier work
    TDSE_HIP(hipEventRecord(b->fork, b->stream));
    for (int s = 0; s < tdse_batch::kSide; ++s) TDSE_HIP(hipStreamWaitEvent(b->side[s], b->fork, 0));
    if ((rc = launch_units(b, nsteps))) return rc;
    // join: the main stream waits for every lane
    for (int s = 0; s < tdse_batch::kSide; ++s) {
        TDSE_HIP(hipEventRecord(b->join[s], b->side[s]));
        TDSE_HIP(hipStreamWaitEvent(b->stream, b->join[s], 0));
    }
    // later work queued on a plan (download, observe, step) waits for the batch
    TDSE_HIP(hipEventRecord(b->done, b->stream));
    for (tdse_plan* p : b->plans) TDSE_HIP(hipStreamWaitEvent(p->stream, b->done, 0));
    return TDSE_OK;
}

int tdse_batch_sync(tdse_batch* b, int timeout_ms) {
    if (!b) return fail(TDSE_E_INVALID, "null batch");
    int rc = set_device(b->ctx);
    if (rc) return rc;
    if (timeout_ms < 0) {
        TDSE_HIP(hipStreamSynchronize(b->stream));
        return TDSE_OK;
    }
    const auto t0 = std::chrono::steady_clock::now();
    for (;;) {
        hipError_t e = hipStreamQuery(b->stream);
        if (e == hipSuccess) return TDSE_OK;
        if (e != hipErrorNotReady) return fail(TDSE_E_DEVICE, hipGetErrorString(e));
        if (std::chrono::duration_cast<std::chrono::milliseconds>(std::chrono::steady_clock::now() - t0)
                .count() >= timeout_ms)
            return fail(TDSE_E_TIMEOUT, "batch wait timed out");
        std::this_thread::yield();
    }
}

int tdse_batch_submit(tdse_ctx* ctx, tdse_plan* const* plans, const int* nsteps, int count) {
    if (!ctx) return fail(TDSE_E_INVALID, "null ctx");
    if (count < 0) return fail(TDSE_E_INVALID, "count < 0");
    if (count > 0 && (!plans || !nsteps)) return fail(TDSE_E_INVALID, "null plans or nsteps");
    // every argument is checked before anything is queued
    for (int i = 0; i < count; ++i) {
        const std::string at = "plan " + std::to_string(i) + ": ";
        if (!plans[i]) return fail(TDSE_E_INVALID, at + "null plan");
        if (plans[i]->ctx != ctx) return fail(TDSE_E_INVALID, at + "belongs to another context");
        if (nsteps[i] < 0) return fail(TDSE_E_INVALID, at + "nsteps < 0");
        if (nsteps[i] > 0 && !plans[i]->has_potential)
            return fail(TDSE_E_STATE, at + "tdse_plan_set_potential not called");
    }
    for (int i = 0; i < count; ++i) {
        const int rc = tdse_step(plans[i], nsteps[i]);
        if (rc) return rc;
    }
    return TDSE_OK;
}

int tdse_sync(tdse_ctx* ctx, int timeout_ms) {
    if (!ctx) return fail(TDSE_E_INVALID, "null ctx");
    int rc = set_device(ctx);
    if (rc) return rc;
    std::lock_guard<std::mutex> lk(ctx->reg_mu);
    if (timeout_ms < 0) {
        for (hipStream_t s : ctx->streams) TDSE_HIP(hipStreamSynchronize(s));
        return TDSE_OK;
    }
    const auto t0 = std::chrono::steady_clock::now();
    for (hipStream_t s : ctx->streams) {
        for (;;) {
            const hipError_t e = hipStreamQuery(s);
            if (e == hipSuccess) break;
            if (e != hipErrorNotReady) return fail(TDSE_E_DEVICE, hipGetErrorString(e));
            if (std::chrono::duration_cast<std::chrono::milliseconds>(
                    std::chrono::steady_clock::now() - t0)
                    .count() >= timeout_ms)
                return fail(TDSE_E_TIMEOUT, "context wait timed out");
            std::this_thread::yield();
        }
    }
    return TDSE_OK;
}

int tdse_batch_stream(tdse_batch* b, void** out) {
    if (!b || !out) return fail(TDSE_E_INVALID, "null argument");
    *out = b->stream;
    return TDSE_OK;
}

int tdse_batch_launches_per_round(const tdse_batch* b, int* out) {
    if (!b || !out) return fail(TDSE_E_INVALID, "null argument");
    int n = 0;
    for (const BatchUnit& u : b->units)
        for (int sub = 0; sub < 4; ++sub) n += static_cast<int>(u.mid[sub].size());
    *out = n;
    return TDSE_OK;
}

}  // extern "C"
