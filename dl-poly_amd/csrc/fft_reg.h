// Compile-time building blocks of the line and plane FFT kernels: exact
// twiddle constants, in-register DFTs of size 2^a 3^b 5^c, and the two stages
// of a line of length N = N1 x N2 split over N2 threads.
#pragma once

#include "fft_common.h"

namespace dlpx {

// ---- exp(2 pi i k / n) at compile time -------------------------------------
// Taylor series on [0, pi/4] after reduction by quadrant and octant; the
// integer part of the reduction is exact, so the values are correct to a few
// ulp (checked against the host's cos/sin by the library tests through the
// transforms they feed).
constexpr double kHalfPiD = 1.57079632679489661923;
constexpr double cx_sin(double x)
{
    double t = x, s = x;
    for (int k = 1; k < 14; k++) {
        t *= -x * x / ((2.0 * k) * (2.0 * k + 1.0));
        s += t;
    }
    return s;
}
constexpr double cx_cos(double x)
{
    double t = 1.0, s = 1.0;
    for (int k = 1; k < 14; k++) {
        t *= -x * x / ((2.0 * k - 1.0) * (2.0 * k));
        s += t;
    }
    return s;
}
struct Cis {
    double c, s;
};
constexpr Cis cx_cis(int k, int n)
{
    k = ((k % n) + n) % n;
    const int q = (4 * k) / n, num = 4 * k - q * n;  // angle = q pi/2 + (pi/2) num/n
    double c = 0.0, s = 0.0;
    if (2 * num <= n) {
        const double a = kHalfPiD * ((double)num / (double)n);
        c = cx_cos(a);
        s = cx_sin(a);
    } else {
        const double b = kHalfPiD * ((double)(n - num) / (double)n);
        c = cx_sin(b);
        s = cx_cos(b);
    }
    switch (q) {
        case 0: return Cis{c, s};
        case 1: return Cis{-s, c};
        case 2: return Cis{-c, -s};
        default: return Cis{s, -c};
    }
}
template <int R>
struct CisTable {
    double c[R], s[R];
    constexpr CisTable() : c(), s()
    {
        for (int k = 0; k < R; k++) {
            const Cis w = cx_cis(k, R);
            c[k] = w.c;
            s[k] = w.s;
        }
    }
};
template <int R>
struct CisOf {
    static constexpr CisTable<R> t{};
};

// ---- in-register DFT of compile-time size -----------------------------------
constexpr int first_factor(int r)
{
    return (r % 4 == 0) ? 4 : (r % 2 == 0) ? 2 : (r % 3 == 0) ? 3 : (r % 5 == 0) ? 5 : r;
}

// In-place DFT of size R, kernel exp(SGN 2 pi i n k / R), on x[OFF + q S]
// (q < R) of a register array.  On return position q holds output index
// perm(q).  R = P Q with P the first factor: Q-point DFTs over n2 of
// x[n1 + P n2], twiddles exp(SGN 2 pi i n1 k2 / R), P-point DFTs over n1.
template <typename T, typename T2, int R, int SGN>
struct RegDft {
    static constexpr bool kBase = R <= 5;
    static constexpr int P = kBase ? R : first_factor(R), Q = R / P;
    static_assert(kBase || (P > 1 && Q > 1), "DFT length must be 2^a 3^b 5^c");
    static constexpr int perm(int q)
    {
        if constexpr (kBase) return q;
        else return Q * (q % P) + RegDft<T, T2, Q, SGN>::perm(q / P);
    }
    template <int OFF, int S>
    __device__ static __forceinline__ void run(T2 *x)
    {
        if constexpr (R == 1) {
            return;
        } else if constexpr (kBase) {
            T2 v[R];
#pragma unroll
            for (int q = 0; q < R; q++) v[q] = x[OFF + q * S];
            dft_small<T, T2, R>(v, SGN);
#pragma unroll
            for (int q = 0; q < R; q++) x[OFF + q * S] = v[q];
        } else {
            run_q<OFF, S, 0>(x);
#pragma unroll
            for (int n1 = 1; n1 < P; n1++)
#pragma unroll
                for (int j = 0; j < Q; j++) {
                    const int e = (n1 * RegDft<T, T2, Q, SGN>::perm(j)) % R;
                    T2 &v = x[OFF + (n1 + P * j) * S];
                    const T2 u = v;
                    if (e == 0) {
                    } else if (4 * e == R) {
                        v = cmuli(u, SGN);
                    } else if (2 * e == R) {
                        v.x = -u.x;
                        v.y = -u.y;
                    } else if (4 * e == 3 * R) {
                        v = cmuli(u, -SGN);
                    } else {
                        const T c = (T)CisOf<R>::t.c[e], s = (T)(SGN * CisOf<R>::t.s[e]);
                        v.x = u.x * c - u.y * s;
                        v.y = u.x * s + u.y * c;
                    }
                }
            run_p<OFF, S, 0>(x);
        }
    }
    // the Q-point DFTs over n2, one per n1 (template recursion keeps every
    // offset a constant)
    template <int OFF, int S, int N1>
    __device__ static __forceinline__ void run_q(T2 *x)
    {
        if constexpr (N1 < P) {
            RegDft<T, T2, Q, SGN>::template run<OFF + N1 * S, S * P>(x);
            run_q<OFF, S, N1 + 1>(x);
        }
    }
    template <int OFF, int S, int J>
    __device__ static __forceinline__ void run_p(T2 *x)
    {
        if constexpr (J < Q) {
            RegDft<T, T2, P, SGN>::template run<OFF + P * J * S, S>(x);
            run_p<OFF, S, J + 1>(x);
        }
    }
};

template <typename T2>
__device__ __forceinline__ T2 zero2()
{
    T2 z;
    z.x = 0;
    z.y = 0;
    return z;
}

// Stage A of one line on thread t: DFT over n1 of x[n1] = element t + N2 n1,
// twiddle by exp(SGN 2 pi i t m1 / N), exchange row ex[(m1 N2 + t) ST] (ST:
// the element stride of the exchange, 1 for a contiguous line).
template <typename T, int N1, int N2, int SGN, int ST = 1>
__device__ __forceinline__ void stage_a(typename C2T<T>::type (&x)[N1], int t, typename C2T<T>::type *ex,
                                        const typename C2T<T>::type *stw)
{
    using T2 = typename C2T<T>::type;
    using D1 = RegDft<T, T2, N1, SGN>;
    D1::template run<0, 1>(x);
#pragma unroll
    for (int p = 0; p < N1; p++) {
        const int m1 = D1::perm(p);
        T2 v = x[p];
        if (m1 != 0) {
            T2 w = stw[t * m1];  // t m1 < N
            if (SGN < 0) w.y = -w.y;
            v = cmul(v, w);
        }
        ex[(m1 * N2 + t) * ST] = v;
    }
}

// Stage B for m1: DFT over t of ex[(m1 N2 + t) ST]; y[p] = X[m1 + N1 perm(p)].
template <typename T, int N1, int N2, int SGN, int ST = 1>
__device__ __forceinline__ void stage_b(int m1, const typename C2T<T>::type *ex, typename C2T<T>::type (&y)[N2])
{
    using T2 = typename C2T<T>::type;
#pragma unroll
    for (int n2 = 0; n2 < N2; n2++) y[n2] = ex[(m1 * N2 + n2) * ST];
    RegDft<T, T2, N2, SGN>::template run<0, 1>(y);
}

// Lengths with a compiled line kernel and their split N = N1 x N2.
#define DLPX_LINE_LENGTHS(X) \
    X(16, 4, 4)              \
    X(20, 5, 4)              \
    X(24, 6, 4)              \
    X(25, 5, 5)              \
    X(27, 9, 3)              \
    X(30, 6, 5)              \
    X(32, 8, 4)              \
    X(36, 6, 6)              \
    X(40, 8, 5)              \
    X(45, 9, 5)              \
    X(48, 12, 4)             \
    X(50, 10, 5)             \
    X(54, 9, 6)              \
    X(60, 12, 5)             \
    X(64, 8, 8)              \
    X(72, 12, 6)             \
    X(80, 16, 5)             \
    X(96, 12, 8)             \
    X(100, 10, 10)           \
    X(120, 12, 10)           \
    X(128, 16, 8)

}  // namespace dlpx
