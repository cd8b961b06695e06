// Definitions shared by the FFT kernels of the SPME convolution: the
// generic Stockham kernels (fft.hip) and the compile-time-length line
// kernels (fft_lines.hip).
#pragma once
#include "common.h"
#include "device_util.h"

namespace dlpx {

struct FftPlan {
    int n, np;
    int rad[24];
    int ls[24];        // Ls, the product of the radices of the passes before p
    float inv_ls[24];  // 1/Ls
};

static FftPlan make_plan(int n)
{
    FftPlan p;
    p.n = n;
    p.np = 0;
    if (n <= 0) return p;  // no mesh (a system without electrostatics)
    int m = n;
    while (m % 4 == 0) { p.rad[p.np++] = 4; m /= 4; }
    while (m % 2 == 0) { p.rad[p.np++] = 2; m /= 2; }
    while (m % 3 == 0) { p.rad[p.np++] = 3; m /= 3; }
    while (m % 5 == 0) { p.rad[p.np++] = 5; m /= 5; }
    int ls = 1;
    for (int q = 0; q < p.np; q++) {
        p.ls[q] = ls;
        p.inv_ls[q] = 1.0f / (float)ls;
        ls *= p.rad[q];
    }
    return p;
}

// x / d and x mod d for 0 <= x < 2^20 from a float reciprocal: the quotient
// estimate is within one of the truth, and one correction step fixes it.
struct FastDiv {
    int d;
    float inv;
};
static FastDiv make_div(int d) { return FastDiv{d, 1.0f / (float)d}; }
__device__ __forceinline__ int fdivmod(int x, FastDiv f, int &r)
{
    int q = (int)((float)x * f.inv);
    r = x - q * f.d;
    if (r < 0) {
        r += f.d;
        q--;
    } else if (r >= f.d) {
        r -= f.d;
        q++;
    }
    return q;
}

template <typename T> struct C2T;
template <> struct C2T<double> { using type = double2; };
template <> struct C2T<float> { using type = float2; };

template <typename T2>
__device__ __forceinline__ T2 cmul(T2 a, T2 b)
{
    T2 r;
    r.x = a.x * b.x - a.y * b.y;
    r.y = a.x * b.y + a.y * b.x;
    return r;
}
template <typename T2> __device__ __forceinline__ T2 cadd(T2 a, T2 b) { T2 r; r.x = a.x + b.x; r.y = a.y + b.y; return r; }
template <typename T2> __device__ __forceinline__ T2 csub(T2 a, T2 b) { T2 r; r.x = a.x - b.x; r.y = a.y - b.y; return r; }
// multiply by s*i
template <typename T2> __device__ __forceinline__ T2 cmuli(T2 a, int s) { T2 r; r.x = -s * a.y; r.y = s * a.x; return r; }

// In-place 4-point DFT with kernel exp(sign 2 pi i jk/4).
template <typename T2>
__device__ __forceinline__ void dft4(T2 &a, T2 &b, T2 &c, T2 &d, int sign)
{
    const T2 s02 = cadd(a, c), d02 = csub(a, c);
    const T2 s13 = cadd(b, d), d13 = cmuli(csub(b, d), sign);
    a = cadd(s02, s13);
    c = csub(s02, s13);
    b = cadd(d02, d13);
    d = csub(d02, d13);
}

// In-place DFT of length R (2, 3, 4, 5) with kernel exp(sign 2 pi i jk/R).
template <typename T, typename T2, int R>
__device__ __forceinline__ void dft_small(T2 *x, int sign)
{
    if (R == 2) {
        T2 a = x[0], b = x[1];
        x[0] = cadd(a, b);
        x[1] = csub(a, b);
    } else if (R == 4) {
        T2 s02 = cadd(x[0], x[2]), d02 = csub(x[0], x[2]);
        T2 s13 = cadd(x[1], x[3]), d13 = cmuli(csub(x[1], x[3]), sign);
        x[0] = cadd(s02, s13);
        x[2] = csub(s02, s13);
        x[1] = cadd(d02, d13);
        x[3] = csub(d02, d13);
    } else if (R == 3) {
        const T h = (T)0.86602540378443864676;  // sqrt(3)/2
        T2 t = cadd(x[1], x[2]), d = csub(x[1], x[2]);
        T2 a;
        a.x = x[0].x - (T)0.5 * t.x;
        a.y = x[0].y - (T)0.5 * t.y;
        T2 b = cmuli(d, sign);
        b.x *= h;
        b.y *= h;
        x[0] = cadd(x[0], t);
        x[1] = cadd(a, b);
        x[2] = csub(a, b);
    } else {  // 5
        const T c1 = (T)0.30901699437494742410, c2 = (T)-0.80901699437494742410;
        const T s1 = (T)0.95105651629515357212, s2 = (T)0.58778525229247312917;
        T2 t1 = cadd(x[1], x[4]), t2 = cadd(x[2], x[3]);
        T2 d1 = csub(x[1], x[4]), d2 = csub(x[2], x[3]);
        T2 a1, a2, p1, p2;
        a1.x = x[0].x + c1 * t1.x + c2 * t2.x;
        a1.y = x[0].y + c1 * t1.y + c2 * t2.y;
        a2.x = x[0].x + c2 * t1.x + c1 * t2.x;
        a2.y = x[0].y + c2 * t1.y + c1 * t2.y;
        // s1 d1 + s2 d2 and s2 d1 - s1 d2, then times sign*i
        p1.x = s1 * d1.x + s2 * d2.x;
        p1.y = s1 * d1.y + s2 * d2.y;
        p2.x = s2 * d1.x - s1 * d2.x;
        p2.y = s2 * d1.y - s1 * d2.y;
        const T2 b1 = cmuli(p1, sign), b2 = cmuli(p2, sign);
        x[0].x = x[0].x + t1.x + t2.x;
        x[0].y = x[0].y + t1.y + t2.y;
        x[1] = cadd(a1, b1);
        x[4] = csub(a1, b1);
        x[2] = cadd(a2, b2);
        x[3] = csub(a2, b2);
    }
}

// Strided lines of the half-spectrum array.  Line L is (a, b) = (L % na_pad,
// L / na_pad); it exists when a < na, starts at a * sa + b * sb and its
// elements are se apart.  na_pad (na rounded up to the lines per workgroup)
// keeps every workgroup's lines on one aligned run of x columns, and the
// array's row stride K1hp (K1/2+1 rounded up to 16) keeps those runs on whole
// 128-byte lines: with rows of K1/2+1 elements every run straddled two.
struct LineMap {
    int nl;        // number of line slots, na_pad * (number of b)
    int na, na_pad;
    long sa, sb;
    long se;
};

__device__ __forceinline__ bool line_ok(const LineMap &lm, int L) { return L < lm.nl && L % lm.na_pad < lm.na; }

template <typename T2>
__device__ __forceinline__ size_t line_base(const LineMap &lm, int L)
{
    return (size_t)(L % lm.na_pad) * lm.sa + (size_t)(L / lm.na_pad) * lm.sb;
}

template <typename T>
struct C2CArgs {
    typename C2T<T>::type *data;
    LineMap lm;
    FftPlan pl;
    const typename C2T<T>::type *tw;
    int sign;
};

template <typename T>
struct ZConvArgs {
    typename C2T<T>::type *data;
    LineMap lm;
    FftPlan pl;
    const typename C2T<T>::type *tw;
    InflParams ip;
    const double *n2;
    const typename C2T<T>::type *theta;  // mode 0: {theta(m), its stress weight} in the layout of data
    double *red;
};

template <typename T>
struct R2CArgs {
    T *Q;
    typename C2T<T>::type *H;
    int nreal_lines, K1h, hs;
    FftPlan pl;
    const typename C2T<T>::type *tw;
    FastDiv div_n, div_h;
    int clear;
};

template <typename T>
struct C2RArgs {
    const typename C2T<T>::type *H;
    T *Q;
    int nreal_lines, K1h, hs;
    FftPlan pl;
    const typename C2T<T>::type *tw;
    FastDiv div_n;
};

// The x and y passes of one z-plane in one workgroup (fft_planes.hip), for
// meshes with K1 = K2 of a line-kernel length: forward (real rows to the
// half spectrum, then the y transforms; the read rows cleared when clear is
// set) and inverse (y transforms, then the half spectrum back to real rows).
template <typename T>
struct PlaneArgs {
    T *Q;                                // real grid, x fastest
    typename C2T<T>::type *H;            // half spectrum, hs elements per row
    int nplanes, hs;
    const typename C2T<T>::type *tw;     // exp(2 pi i k / K1), k < K1
    int clear;
};

constexpr int kConvRedBytes = 6 * (512 / 64) * sizeof(double);

// Line kernels of compile-time length (fft_lines.hip) for the lengths
// lines_supported() accepts; lines_lpb() lines (or x-line pairs) per
// workgroup, lines_lds() their dynamic LDS (kind 0 y pass, 1 z pass, 2 forward
// x, 3 inverse x).
bool lines_supported(int n);
template <typename T> size_t plane_lds(int n);
template <typename T> int plane_run(Ctx *c, hipStream_t s, bool inverse, int n, const PlaneArgs<T> &a);
template <typename T> int plane_batch(Ctx *c, hipStream_t s, bool inverse, int n, const BatchLaunch &bl);
template <typename T> int lines_lpb();
template <typename T> size_t lines_lds(int kind, int n);
template <typename T> int lines_c2c(Ctx *c, hipStream_t s, const C2CArgs<T> &a, int nb);
template <typename T> int lines_zconv(Ctx *c, hipStream_t s, const ZConvArgs<T> &a, int nb);
template <typename T> int lines_x(Ctx *c, hipStream_t s, const R2CArgs<T> *fwd, const C2RArgs<T> *inv, int nb);
// one batched launch of a line-kernel group (FftGroup kind and length)
template <typename T> int lines_batch(Ctx *c, hipStream_t s, int kind, int n, const BatchLaunch &bl);

}  // namespace dlpx
