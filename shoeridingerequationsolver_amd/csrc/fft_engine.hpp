// fft_engine.hpp -- device-side building blocks of the line FFT.
//
// A line of N complex points is transformed by one group of T = N / E threads;
// every thread keeps E points in registers.  The transform is a Stockham
// autosort FFT (natural order in and out) over a compile-time list of radices
// whose product is N; every radix divides E, so a thread runs E / R butterflies
// per pass.  Between passes the points are exchanged through LDS.
//
// Register pattern of a pass with radix R (j = t + b*T, b < E/R):
//     v[b*R + r] = x[j + r*(N/R)]
// Output index of that butterfly's r-th result (Ns = product of earlier radices):
//     o = (j / Ns) * Ns * R + (j % Ns) + r * Ns
// After the last pass (Ns = N / R_last) the register pattern is
// v[b*R_last + r] = X[j + r*(N/R_last)], which is exactly the input pattern of
// a first pass with the same radix.  That is why an inverse transform whose
// radix list is the forward list reversed can start from the forward
// transform's registers without an exchange, and vice versa.
//
// Twiddles between passes come from a table W[m] = exp(-2 pi i m / N),
// m < N, built on the host in FP64 (rounded to the path precision); the
// inverse uses the conjugate.  Codelet constants are evaluated at compile time
// with an exact rational octant reduction and a Taylor series in double.
#pragma once

#include <hip/hip_runtime.h>

#include <type_traits>

namespace tdse {

template <typename T>
struct alignas(2 * sizeof(T)) cx {
    T x, y;
};

template <typename T>
__device__ __forceinline__ cx<T> operator+(cx<T> a, cx<T> b) { return {a.x + b.x, a.y + b.y}; }
template <typename T>
__device__ __forceinline__ cx<T> operator-(cx<T> a, cx<T> b) { return {a.x - b.x, a.y - b.y}; }
template <typename T>
__device__ __forceinline__ cx<T> cmul(cx<T> a, cx<T> b) {
    return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
template <typename T>
__device__ __forceinline__ cx<T> cscale(cx<T> a, T s) { return {a.x * s, a.y * s}; }
template <typename T>
__device__ __forceinline__ cx<T> conj(cx<T> a) { return {a.x, -a.y}; }
// multiply by DIR*i: DIR = -1 -> -i, DIR = +1 -> +i
template <int DIR, typename T>
__device__ __forceinline__ cx<T> mul_i(cx<T> a) {
    if constexpr (DIR < 0) return {a.y, -a.x};
    else return {-a.y, a.x};
}

// x * K for a compile-time constant K.  In FP32 the constant is split into
// hi + lo (hi = K rounded to float, lo = the rounding error) so the product
// carries K to about 48 bits: a rounded constant multiplied into the same
// fraction of the state every step would otherwise bias its norm coherently.
__device__ __forceinline__ double mulk(double x, double k) { return x * k; }
__device__ __forceinline__ float mulk(float x, double k) {
    const float hi = static_cast<float>(k);
    const float lo = static_cast<float>(k - static_cast<double>(hi));
    return fmaf(x, hi, x * lo);
}

// ---- compile-time sin/cos of 2*pi*m/R --------------------------------------

constexpr double kPi = 3.14159265358979323846;

constexpr double taylor_sin(double x) {
    double term = x, sum = x;
    for (int k = 1; k < 14; ++k) {
        term *= -x * x / ((2.0 * k) * (2.0 * k + 1.0));
        sum += term;
    }
    return sum;
}
constexpr double taylor_cos(double x) {
    double term = 1.0, sum = 1.0;
    for (int k = 1; k < 14; ++k) {
        term *= -x * x / ((2.0 * k - 1.0) * (2.0 * k));
        sum += term;
    }
    return sum;
}

struct SinCos {
    double c, s;
};

// cos/sin of 2*pi*num/den, reduced exactly on the rational to [0, pi/4].
constexpr SinCos sincos_2pi(long long num, long long den) {
    num %= den;
    if (num < 0) num += den;
    // angle in [0, 2pi): reflect the lower half plane
    bool neg_s = false;
    if (2 * num > den) {
        num = den - num;
        neg_s = true;
    }
    // now angle = 2 pi num/den in [0, pi]; reflect about pi/2
    bool neg_c = false;
    long long n2 = num, d2 = den;
    if (4 * num > den) {  // angle > pi/2: cos(a) = -cos(pi - a), sin(a) = sin(pi - a)
        n2 = den - 2 * num;  // (pi - a) = 2 pi (den - 2 num) / (2 den)
        d2 = 2 * den;
        neg_c = true;
    }
    // angle in [0, pi/2]; swap about pi/4
    bool swap = false;
    if (8 * n2 > d2) {  // a > pi/4: use pi/2 - a = 2 pi (d2 - 4 n2) / (4 d2)
        long long n3 = d2 - 4 * n2;
        d2 = 4 * d2;
        n2 = n3;
        swap = true;
    }
    const double x = 2.0 * kPi * static_cast<double>(n2) / static_cast<double>(d2);
    double c = taylor_cos(x), s = taylor_sin(x);
    if (swap) {
        const double t = c;
        c = s;
        s = t;
    }
    if (neg_c) c = -c;
    if (neg_s) s = -s;
    return SinCos{c, s};
}


// a * W_R^m with sign DIR: a * exp(DIR * 2 pi i m / R), constant folded.
// (A fused form for FP32 -- one fma chain per component, 8 operations instead
// of 10, and one compensated product per component at odd multiples of pi/4
// -- measured 2-3% faster on the 256^3 x and z sweeps but moved the FP32 norm
// of a 96^3 solve by -1.9e-6 over 500 steps: not kept.)
template <typename T, int DIR, int M, int R>
__device__ __forceinline__ cx<T> mul_cw(cx<T> a) {
    constexpr SinCos sc = sincos_2pi(M, R);
    constexpr double c = sc.c, s = DIR * sc.s;
    return {mulk(a.x, c) - mulk(a.y, s), mulk(a.x, s) + mulk(a.y, c)};
}

// a * W_R^m at an odd multiple of pi/4 (|cos| = |sin| = 1/sqrt 2) in FP32:
// ((+-x -+ y) + i (+-x +- y)) / sqrt 2, the sums rounded once and the
// constant applied compensated (mulk): 6 operations instead of 10, one
// rounding fewer per component.
template <typename T, int DIR, int M, int R>
__device__ __forceinline__ cx<T> mul_octant(cx<T> a) {
    constexpr SinCos sc = sincos_2pi(M, R);
    constexpr bool cpos = sc.c > 0, spos = DIR * sc.s > 0;
    constexpr double h = 0.70710678118654752440;
    const T u = cpos ? (spos ? a.x - a.y : a.x + a.y) : (spos ? -a.x - a.y : a.y - a.x);
    const T w = cpos ? (spos ? a.x + a.y : a.y - a.x) : (spos ? a.x - a.y : -a.x - a.y);
    return {mulk(u, h), mulk(w, h)};
}

template <typename T, int DIR, int M, int R>
__device__ __forceinline__ cx<T> twiddle_const(cx<T> a) {
    constexpr int m = ((M % R) + R) % R;
    if constexpr (m == 0) {
        return a;
    } else if constexpr (4 * m == R) {
        return mul_i<DIR>(a);
    } else if constexpr (2 * m == R) {
        return cx<T>{-a.x, -a.y};
    } else if constexpr (4 * m == 3 * R) {
        return mul_i<-DIR>(a);
    } else if constexpr (sizeof(T) == 4 && (8 * m) % R == 0) {
        return mul_octant<T, DIR, m, R>(a);
    } else {
        return mul_cw<T, DIR, m, R>(a);
    }
}

// ---- small DFT codelets (in place on v[0..R-1], natural order) --------------

template <int R, int DIR>
struct Dft;

template <int DIR>
struct Dft<1, DIR> {
    template <typename T>
    __device__ __forceinline__ static void run(cx<T>*) {}
};

template <int DIR>
struct Dft<2, DIR> {
    template <typename T>
    __device__ __forceinline__ static void run(cx<T>* v) {
        const cx<T> a = v[0], b = v[1];
        v[0] = a + b;
        v[1] = a - b;
    }
};

template <int DIR>
struct Dft<3, DIR> {
    template <typename T>
    __device__ __forceinline__ static void run(cx<T>* v) {
        constexpr double s = DIR * 0.86602540378443864676;
        const cx<T> t1 = v[1] + v[2], t2 = v[1] - v[2];
        const cx<T> m = v[0] - cscale(t1, static_cast<T>(0.5));
        const cx<T> is = {-mulk(t2.y, s), mulk(t2.x, s)};  // i*s*t2
        v[0] = v[0] + t1;
        v[1] = m + is;
        v[2] = m - is;
    }
};

template <int DIR>
struct Dft<4, DIR> {
    template <typename T>
    __device__ __forceinline__ static void run(cx<T>* v) {
        const cx<T> a = v[0] + v[2], b = v[0] - v[2];
        const cx<T> c = v[1] + v[3], d = mul_i<DIR>(v[1] - v[3]);
        v[0] = a + c;
        v[2] = a - c;
        v[1] = b + d;
        v[3] = b - d;
    }
};

template <int DIR>
struct Dft<5, DIR> {
    template <typename T>
    __device__ __forceinline__ static void run(cx<T>* v) {
        constexpr double c1 = 0.30901699437494742410;   // cos(2pi/5)
        constexpr double c2 = -0.80901699437494742410;  // cos(4pi/5)
        constexpr double s1 = DIR * 0.95105651629515357212;
        constexpr double s2 = DIR * 0.58778525229247312917;
        const cx<T> t1 = v[1] + v[4], t2 = v[2] + v[3];
        const cx<T> t3 = v[1] - v[4], t4 = v[2] - v[3];
        const cx<T> m1 = {v[0].x + mulk(t1.x, c1) + mulk(t2.x, c2),
                          v[0].y + mulk(t1.y, c1) + mulk(t2.y, c2)};
        const cx<T> m2 = {v[0].x + mulk(t1.x, c2) + mulk(t2.x, c1),
                          v[0].y + mulk(t1.y, c2) + mulk(t2.y, c1)};
        const cx<T> n1 = {mulk(t3.x, s1) + mulk(t4.x, s2), mulk(t3.y, s1) + mulk(t4.y, s2)};
        const cx<T> n2 = {mulk(t3.x, s2) - mulk(t4.x, s1), mulk(t3.y, s2) - mulk(t4.y, s1)};
        v[0] = v[0] + t1 + t2;
        v[1] = {m1.x - n1.y, m1.y + n1.x};
        v[4] = {m1.x + n1.y, m1.y - n1.x};
        v[2] = {m2.x - n2.y, m2.y + n2.x};
        v[3] = {m2.x + n2.y, m2.y - n2.x};
    }
};

// Composite R = A*B (Cooley-Tukey in registers):
//   Y[n1][k2] = DFT_B over n2 of x[A*n2 + n1],  Y *= W_R^{n1*k2},
//   X[k2 + B*k1] = DFT_A over n1 of Y[n1][k2].
template <int A, int B, int DIR>
struct DftCT {
    template <typename T>
    __device__ __forceinline__ static void run(cx<T>* v) {
        constexpr int R = A * B;
        cx<T> y[R];  // y[n1*B + k2]
#pragma unroll
        for (int n1 = 0; n1 < A; ++n1) {
            cx<T> t[B];
#pragma unroll
            for (int n2 = 0; n2 < B; ++n2) t[n2] = v[A * n2 + n1];
            Dft<B, DIR>::run(t);
#pragma unroll
            for (int k2 = 0; k2 < B; ++k2) y[n1 * B + k2] = t[k2];
        }
        apply_tw<T>(y, std::integral_constant<int, 0>{});
#pragma unroll
        for (int k2 = 0; k2 < B; ++k2) {
            cx<T> t[A];
#pragma unroll
            for (int n1 = 0; n1 < A; ++n1) t[n1] = y[n1 * B + k2];
            Dft<A, DIR>::run(t);
#pragma unroll
            for (int k1 = 0; k1 < A; ++k1) v[k2 + B * k1] = t[k1];
        }
    }

    // compile-time loop over (n1, k2) so every twiddle is a constant
    template <typename T, int I>
    __device__ __forceinline__ static void apply_tw(cx<T>* y, std::integral_constant<int, I>) {
        if constexpr (I < A * B) {
            constexpr int n1 = I / B, k2 = I % B;
            y[I] = twiddle_const<T, DIR, n1 * k2, A * B>(y[I]);
            apply_tw<T>(y, std::integral_constant<int, I + 1>{});
        }
    }
};

template <int DIR> struct Dft<6, DIR> : DftCT<2, 3, DIR> {};
template <int DIR> struct Dft<8, DIR> : DftCT<2, 4, DIR> {};
template <int DIR> struct Dft<10, DIR> : DftCT<2, 5, DIR> {};
template <int DIR> struct Dft<12, DIR> : DftCT<4, 3, DIR> {};
template <int DIR> struct Dft<16, DIR> : DftCT<4, 4, DIR> {};
template <int DIR> struct Dft<32, DIR> : DftCT<4, 8, DIR> {};

// ---- radix lists -------------------------------------------------------------

template <int... Rs>
struct Radices {
    static constexpr int count = sizeof...(Rs);
    static constexpr int product = (1 * ... * Rs);
};

template <class L, int... Acc>
struct ReverseImpl;
template <int... Acc>
struct ReverseImpl<Radices<>, Acc...> {
    using type = Radices<Acc...>;
};
template <int R, int... Rest, int... Acc>
struct ReverseImpl<Radices<R, Rest...>, Acc...> {
    using type = typename ReverseImpl<Radices<Rest...>, R, Acc...>::type;
};
template <class L>
using Reverse = typename ReverseImpl<L>::type;

template <class L>
struct First;
template <int R, int... Rest>
struct First<Radices<R, Rest...>> {
    static constexpr int value = R;
};
template <class L>
struct Last;
template <int R>
struct Last<Radices<R>> {
    static constexpr int value = R;
};
template <int R, int R2, int... Rest>
struct Last<Radices<R, R2, Rest...>> {
    static constexpr int value = Last<Radices<R2, Rest...>>::value;
};

template <class L>
struct MinRadix;
template <int R>
struct MinRadix<Radices<R>> {
    static constexpr int value = R;
};
template <int R, int R2, int... Rest>
struct MinRadix<Radices<R, R2, Rest...>> {
    static constexpr int value =
        R < MinRadix<Radices<R2, Rest...>>::value ? R : MinRadix<Radices<R2, Rest...>>::value;
};

// ---- FP64 phase application ------------------------------------------------------

// a * k with a widened to FP64 and the product rounded once to T.
template <typename T>
__device__ __forceinline__ cx<T> mul_d(cx<T> a, cx<double> k) {
    const double x = static_cast<double>(a.x), y = static_cast<double>(a.y);
    return {static_cast<T>(x * k.x - y * k.y), static_cast<T>(x * k.y + y * k.x)};
}

// ---- one Stockham pass in registers -------------------------------------------

// Twiddle then DFT for every butterfly this thread owns.  `tw` holds
// W_N^m = exp(-2 pi i m / N) in FP64 (forward sign).  The butterfly's
// twiddles W^{r k} (r = 1..R-1) are applied in FP64, so no float-rounded
// twiddle ever scales the state.  With TABLE, `tw` holds all N entries and
// each twiddle is read from it (r k step < N); otherwise `tw` holds the
// first N / (least radix) and the twiddles are generated in FP64 from the
// one entry W^k by repeated multiplication (relative error below 2e-15 for
// R <= 16).
template <typename T, int N, int E, int R, int NS, int DIR, bool TABLE = false>
__device__ __forceinline__ void pass_compute(cx<T> (&v)[E], int t,
                                             const cx<double>* __restrict__ tw) {
    constexpr int TT = N / E;
#pragma unroll
    for (int b = 0; b < E / R; ++b) {
        if constexpr (NS > 1 && TABLE) {
            // W^{r k step}: the offset is stepped through an empty asm so the
            // compiler forms each read's address in turn instead of holding
            // R - 1 offsets live
            const int j = t + b * TT;
            const int ks = (j % NS) * (N / (NS * R));
            int m = 0;
#pragma unroll
            for (int r = 1; r < R; ++r) {
                m += ks;
                asm volatile("" : "+v"(m));
                cx<double> w = tw[m];
                if constexpr (DIR > 0) w.y = -w.y;
                v[b * R + r] = mul_d(v[b * R + r], w);
                __builtin_amdgcn_sched_barrier(0);
            }
        } else if constexpr (NS > 1) {
            const int j = t + b * TT;
            const int k = j % NS;
            constexpr int step = N / (NS * R);
            cx<double> w1 = tw[k * step];
            if constexpr (DIR > 0) w1.y = -w1.y;
            cx<double> w = w1;
#pragma unroll
            for (int r = 1; r < R; ++r) {
                v[b * R + r] = mul_d(v[b * R + r], w);
                if (r + 1 < R) w = cmul(w, w1);
                // apply each twiddle as it is formed: left free, the
                // scheduler forms all R - 1 first (FP64, 4 VGPRs each) and
                // the strided two-transform sweeps drop to 2 waves per SIMD
                __builtin_amdgcn_sched_barrier(0);
            }
        }
        Dft<R, DIR>::run(&v[b * R]);
    }
}

// Store pass outputs through `put(index, value)` (Stockham output order).
template <typename T, int N, int E, int R, int NS, class Put>
__device__ __forceinline__ void pass_scatter(const cx<T> (&v)[E], int t, Put&& put) {
    constexpr int TT = N / E;
#pragma unroll
    for (int b = 0; b < E / R; ++b) {
        const int j = t + b * TT;
        const int base = (j / NS) * NS * R + (j % NS);
#pragma unroll
        for (int r = 0; r < R; ++r) put(base + r * NS, v[b * R + r]);
    }
}

// Load the input pattern of a pass with radix R through `get(index)`.
template <typename T, int N, int E, int R, class Get>
__device__ __forceinline__ void pass_gather(cx<T> (&v)[E], int t, Get&& get) {
    constexpr int TT = N / E;
#pragma unroll
    for (int b = 0; b < E / R; ++b) {
        const int j = t + b * TT;
#pragma unroll
        for (int r = 0; r < R; ++r) v[b * R + r] = get(j + r * (N / R));
    }
}

// Visit the points a thread holds in the register pattern of radix R:
// fn(register slot, line index of that point, j of its butterfly, r).
template <int N, int E, int R, class Fn>
__device__ __forceinline__ void for_each_point(int t, Fn&& fn) {
    constexpr int TT = N / E;
#pragma unroll
    for (int b = 0; b < E / R; ++b) {
        const int j = t + b * TT;
#pragma unroll
        for (int r = 0; r < R; ++r) fn(b * R + r, j + r * (N / R), j, r * (N / R));
    }
}

// Full transform: passes over the radix list with LDS exchanges in between.
// `x` is the exchange buffer: x.sync(), x.put(i, v), x.get(i); X::kTwTable
// says whether `tw` is the full twiddle table (see pass_compute).
template <typename T, int N, int E, int DIR, int NS, class L>
struct Transform;

template <typename T, int N, int E, int DIR, int NS, int R>
struct Transform<T, N, E, DIR, NS, Radices<R>> {
    template <class X>
    __device__ __forceinline__ static void run(cx<T> (&v)[E], int t, const cx<double>* tw, X& x) {
        pass_compute<T, N, E, R, NS, DIR, X::kTwTable>(v, t, tw);
    }
};

template <typename T, int N, int E, int DIR, int NS, int R, int R2, int... Rest>
struct Transform<T, N, E, DIR, NS, Radices<R, R2, Rest...>> {
    template <class X>
    __device__ __forceinline__ static void run(cx<T> (&v)[E], int t, const cx<double>* tw, X& x) {
        pass_compute<T, N, E, R, NS, DIR, X::kTwTable>(v, t, tw);
        x.sync();
        pass_scatter<T, N, E, R, NS>(v, t, [&](int i, cx<T> z) { x.put(i, z); });
        x.sync();
        pass_gather<T, N, E, R2>(v, t, [&](int i) { return x.get(i); });
        Transform<T, N, E, DIR, NS * R, Radices<R2, Rest...>>::run(v, t, tw, x);
    }
};

}  // namespace tdse
