// 3D FFT of the SPME grid: batched 1D Stockham transforms in LDS for
// 2^a 3^b 5^c lengths (the sizes parallel_fft.F90:1999-2054 accepts),
// replacing DL_POLY's GPFA/DaFT (gpfa235.F90).  Convention of
// gpfa235.F90:72: sign +1 is sum_k c_k exp(+2 pi i j k/n), no normalisation.
//
// The charge grid is real, so the x axis is done as a real-to-complex
// transform: two real x-lines are packed as one complex line, transformed,
// and split into their half spectra (k = 0..K1/2).  The y and z passes then
// run over the K1/2+1 remaining columns only; the inverse mirrors this and
// writes the real potential grid back over the charge grid.
#include <vector>

#include "common.h"

namespace dlpx {

struct FftPlan {
    int n, np;
    int rad[24];
};

static FftPlan make_plan(int n)
{
    FftPlan p;
    p.n = n;
    p.np = 0;
    int m = n;
    while (m % 4 == 0) { p.rad[p.np++] = 4; m /= 4; }
    while (m % 2 == 0) { p.rad[p.np++] = 2; m /= 2; }
    while (m % 3 == 0) { p.rad[p.np++] = 3; m /= 3; }
    while (m % 5 == 0) { p.rad[p.np++] = 5; m /= 5; }
    return p;
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

// One Stockham pass of radix R: butterfly j of line ll reads
// src[j + q m] (m = n/R), twiddles by exp(sign 2 pi i q k/(Ls R)) with
// k = j mod Ls, and writes dst[(j - k) R + k + s Ls].
template <typename T, typename T2, int R>
__device__ __forceinline__ void stockham_pass(const T2 *src, T2 *dst, int nl, int ns, int n, int Ls,
                                              const T2 *__restrict__ tw, int sign)
{
    const int m = n / R, tstride = n / (Ls * R);
    for (int b = threadIdx.x; b < nl * m; b += blockDim.x) {
        const int ll = b / m, j = b - ll * m;
        const int k = j % Ls;
        T2 x[R];
#pragma unroll
        for (int q = 0; q < R; q++) x[q] = src[ll * ns + j + q * m];
#pragma unroll
        for (int q = 1; q < R; q++) {
            T2 w = tw[q * k * tstride];
            w.y = sign > 0 ? w.y : -w.y;
            x[q] = cmul(x[q], w);
        }
        dft_small<T, T2, R>(x, sign);
        const int base = (j - k) * R + k;
#pragma unroll
        for (int s = 0; s < R; s++) dst[ll * ns + base + s * Ls] = x[s];
    }
}

// Stockham passes over nl lines of length n stored at buf0[ll*ns + e]; the
// result ends in the returned buffer.
template <typename T, typename T2>
__device__ T2 *stockham(T2 *buf0, T2 *buf1, int nl, int ns, const FftPlan &pl, const T2 *__restrict__ tw, int sign)
{
    const int n = pl.n;
    T2 *src = buf0, *dst = buf1;
    int Ls = 1;
    for (int p = 0; p < pl.np; p++) {
        const int r = pl.rad[p];
        if (r == 4) stockham_pass<T, T2, 4>(src, dst, nl, ns, n, Ls, tw, sign);
        else if (r == 2) stockham_pass<T, T2, 2>(src, dst, nl, ns, n, Ls, tw, sign);
        else if (r == 3) stockham_pass<T, T2, 3>(src, dst, nl, ns, n, Ls, tw, sign);
        else stockham_pass<T, T2, 5>(src, dst, nl, ns, n, Ls, tw, sign);
        __syncthreads();
        T2 *t = src;
        src = dst;
        dst = t;
        Ls *= r;
    }
    return src;
}

struct LineMap {
    int n;         // transform length
    int nl;        // number of lines
    int lpb;       // lines per block
    long se;       // element stride
    long sa, sb;   // line L -> (L % na) * sa + (L / na) * sb
    int na;
};

// Complex-to-complex transforms of strided lines, in place.
template <typename T>
__global__ void __launch_bounds__(128) k_fft_c2c(typename C2T<T>::type *data, LineMap lm, FftPlan pl,
                                                 const typename C2T<T>::type *__restrict__ tw, int sign)
{
    using T2 = typename C2T<T>::type;
    extern __shared__ __align__(16) unsigned char smem[];
    const int n = lm.n, ns = n + 1;
    T2 *b0 = (T2 *)smem, *b1 = b0 + lm.lpb * ns;
    const int L0 = blockIdx.x * lm.lpb;
    const int nl = min(lm.lpb, lm.nl - L0);
    for (int idx = threadIdx.x; idx < lm.lpb * n; idx += blockDim.x) {
        const int ll = idx % lm.lpb, e = idx / lm.lpb;
        if (ll < nl) {
            const int L = L0 + ll;
            b0[ll * ns + e] = data[(L % lm.na) * lm.sa + (long)(L / lm.na) * lm.sb + e * lm.se];
        }
    }
    __syncthreads();
    T2 *res = stockham<T, T2>(b0, b1, nl, ns, pl, tw, sign);
    for (int idx = threadIdx.x; idx < lm.lpb * n; idx += blockDim.x) {
        const int ll = idx % lm.lpb, e = idx / lm.lpb;
        if (ll < nl) {
            const int L = L0 + ll;
            data[(L % lm.na) * lm.sa + (long)(L / lm.na) * lm.sb + e * lm.se] = res[ll * ns + e];
        }
    }
}

// x axis, forward: real lines 2P and 2P+1 of Q (K1 contiguous values each)
// packed as one complex line; output half spectra to H (K1h per line).
template <typename T>
__global__ void __launch_bounds__(128) k_fft_x_r2c(const T *__restrict__ Q, typename C2T<T>::type *H, int nreal_lines,
                                                   int K1h, int lpb, FftPlan pl,
                                                   const typename C2T<T>::type *__restrict__ tw)
{
    using T2 = typename C2T<T>::type;
    extern __shared__ __align__(16) unsigned char smem[];
    const int n = pl.n, ns = n + 1;
    T2 *b0 = (T2 *)smem, *b1 = b0 + lpb * ns;
    const int npairs = (nreal_lines + 1) / 2;
    const int P0 = blockIdx.x * lpb;
    const int np = min(lpb, npairs - P0);
    for (int idx = threadIdx.x; idx < lpb * n; idx += blockDim.x) {
        const int ll = idx / n, e = idx - ll * n;
        if (ll < np) {
            const int a = 2 * (P0 + ll), b = a + 1;
            T2 z;
            z.x = Q[(size_t)a * n + e];
            z.y = b < nreal_lines ? Q[(size_t)b * n + e] : (T)0;
            b0[ll * ns + e] = z;
        }
    }
    __syncthreads();
    T2 *res = stockham<T, T2>(b0, b1, np, ns, pl, tw, +1);
    for (int idx = threadIdx.x; idx < lpb * K1h; idx += blockDim.x) {
        const int ll = idx / K1h, k = idx - ll * K1h;
        if (ll < np) {
            const int a = 2 * (P0 + ll), b = a + 1;
            const T2 z = res[ll * ns + k];
            const T2 zc = res[ll * ns + (k == 0 ? 0 : n - k)];
            T2 A, B;
            A.x = (T)0.5 * (z.x + zc.x);
            A.y = (T)0.5 * (z.y - zc.y);
            B.x = (T)0.5 * (z.y + zc.y);
            B.y = (T)-0.5 * (z.x - zc.x);
            H[(size_t)a * K1h + k] = A;
            if (b < nreal_lines) H[(size_t)b * K1h + k] = B;
        }
    }
}

// x axis, inverse: rebuild the full Hermitian line of two half spectra,
// transform with sign -1 and write the real parts to Q.
template <typename T>
__global__ void __launch_bounds__(128) k_fft_x_c2r(const typename C2T<T>::type *__restrict__ H, T *Q,
                                                   int nreal_lines, int K1h, int lpb, FftPlan pl,
                                                   const typename C2T<T>::type *__restrict__ tw)
{
    using T2 = typename C2T<T>::type;
    extern __shared__ __align__(16) unsigned char smem[];
    const int n = pl.n, ns = n + 1;
    T2 *b0 = (T2 *)smem, *b1 = b0 + lpb * ns;
    const int npairs = (nreal_lines + 1) / 2;
    const int P0 = blockIdx.x * lpb;
    const int np = min(lpb, npairs - P0);
    for (int idx = threadIdx.x; idx < lpb * n; idx += blockDim.x) {
        const int ll = idx / n, k = idx - ll * n;
        if (ll < np) {
            const int a = 2 * (P0 + ll), b = a + 1;
            const bool hb = b < nreal_lines;
            T2 z;
            if (k < K1h) {
                const T2 A = H[(size_t)a * K1h + k];
                T2 B;
                B.x = hb ? H[(size_t)b * K1h + k].x : (T)0;
                B.y = hb ? H[(size_t)b * K1h + k].y : (T)0;
                z.x = A.x - B.y;
                z.y = A.y + B.x;
            } else {
                const int km = n - k;
                const T2 A = H[(size_t)a * K1h + km];
                T2 B;
                B.x = hb ? H[(size_t)b * K1h + km].x : (T)0;
                B.y = hb ? H[(size_t)b * K1h + km].y : (T)0;
                z.x = A.x + B.y;
                z.y = -A.y + B.x;
            }
            b0[ll * ns + k] = z;
        }
    }
    __syncthreads();
    T2 *res = stockham<T, T2>(b0, b1, np, ns, pl, tw, -1);
    for (int idx = threadIdx.x; idx < lpb * n; idx += blockDim.x) {
        const int ll = idx / n, e = idx - ll * n;
        if (ll < np) {
            const int a = 2 * (P0 + ll), b = a + 1;
            const T2 z = res[ll * ns + e];
            Q[(size_t)a * n + e] = z.x;
            if (b < nreal_lines) Q[(size_t)b * n + e] = z.y;
        }
    }
}

constexpr int kFftThreads = 128;

template <typename T>
static int lines_per_block(int n)
{
    return std::max(1, 1024 / (n + 1));
}

template <typename T>
static size_t lds_bytes(int lpb, int n)
{
    return 2 * sizeof(typename C2T<T>::type) * (size_t)lpb * (n + 1);
}

template <typename T>
static int run_c2c(Ctx *c, int axis, int sign)
{
    using T2 = typename C2T<T>::type;
    LineMap lm;
    const int K1h = c->K1h, K2 = c->K[1], K3 = c->K[2];
    if (axis == 1) {
        lm.n = K2; lm.nl = K1h * K3; lm.se = K1h; lm.na = K1h; lm.sa = 1; lm.sb = (long)K1h * K2;
    } else {
        lm.n = K3; lm.nl = K1h * K2; lm.se = (long)K1h * K2; lm.na = K1h * K2; lm.sa = 1; lm.sb = 0;
    }
    lm.lpb = lines_per_block<T>(lm.n);
    const FftPlan pl = make_plan(lm.n);
    const T2 *tw = (const T2 *)c->tw + (axis == 1 ? c->K[0] : c->K[0] + c->K[1]);
    const size_t lds = lds_bytes<T>(lm.lpb, lm.n);
    DLPX_HIP(c, hipFuncSetAttribute((const void *)k_fft_c2c<T>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    hipLaunchKernelGGL(k_fft_c2c<T>, dim3(ceil_div(lm.nl, lm.lpb)), dim3(kFftThreads), lds, c->stream, (T2 *)c->H, lm, pl, tw,
                       sign);
    DLPX_HIP(c, hipGetLastError());
    return 0;
}

template <typename T>
static int run_x(Ctx *c, bool forward)
{
    using T2 = typename C2T<T>::type;
    const int n = c->K[0], nlines = c->K[1] * c->K[2];
    const int lpb = lines_per_block<T>(n);
    const int npairs = (nlines + 1) / 2;
    const FftPlan pl = make_plan(n);
    const size_t lds = lds_bytes<T>(lpb, n);
    if (forward) {
        DLPX_HIP(c, hipFuncSetAttribute((const void *)k_fft_x_r2c<T>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
        hipLaunchKernelGGL(k_fft_x_r2c<T>, dim3(ceil_div(npairs, lpb)), dim3(kFftThreads), lds, c->stream, (const T *)c->Q,
                           (T2 *)c->H, nlines, c->K1h, lpb, pl, (const T2 *)c->tw);
    } else {
        DLPX_HIP(c, hipFuncSetAttribute((const void *)k_fft_x_c2r<T>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
        hipLaunchKernelGGL(k_fft_x_c2r<T>, dim3(ceil_div(npairs, lpb)), dim3(kFftThreads), lds, c->stream, (const T2 *)c->H,
                           (T *)c->Q, nlines, c->K1h, lpb, pl, (const T2 *)c->tw);
    }
    DLPX_HIP(c, hipGetLastError());
    return 0;
}

template <typename T>
static int forward_t(Ctx *c)
{
    int r;
    if ((r = run_x<T>(c, true))) return r;
    if ((r = run_c2c<T>(c, 1, +1))) return r;
    return run_c2c<T>(c, 2, +1);
}

template <typename T>
static int inverse_t(Ctx *c)
{
    int r;
    if ((r = run_c2c<T>(c, 2, -1))) return r;
    if ((r = run_c2c<T>(c, 1, -1))) return r;
    return run_x<T>(c, false);
}

int fft_forward(Ctx *c) { return c->prm.precision == DLPX_MIXED ? forward_t<float>(c) : forward_t<double>(c); }
int fft_inverse(Ctx *c) { return c->prm.precision == DLPX_MIXED ? inverse_t<float>(c) : inverse_t<double>(c); }

}  // namespace dlpx
