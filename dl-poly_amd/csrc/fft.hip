// 3D FFT of the SPME grid: batched 1D Stockham transforms in LDS for
// 2^a 3^b 5^c lengths (the sizes parallel_fft.F90:1999-2054 accepts), radix
// 4, 2, 3 and 5 passes,
// replacing DL_POLY's GPFA/DaFT (gpfa235.F90).  Convention of
// gpfa235.F90:72: sign +1 is sum_k c_k exp(+2 pi i j k/n), no normalisation.
//
// The charge grid is real, so the x axis is done as a real-to-complex
// transform: two real x-lines are packed as one complex line, transformed,
// and split into their half spectra (k = 0..K1/2).  The y and z passes then
// run over the K1/2+1 remaining columns only; the inverse mirrors this and
// writes the real potential grid back over the charge grid.
#include <algorithm>
#include <vector>

#include "fft_common.h"

namespace dlpx {

// 512 threads per workgroup: the y and z passes have few workgroups (one per
// 16 lines) and are latency-bound, so each needs many wavefronts in flight.
constexpr int kFftThreads = 512;

// One Stockham pass of radix R over one line: butterfly j (0 <= j < n/R)
// reads src[j + q m] (m = n/R), twiddles input q by exp(sign 2 pi i q k/(Ls R))
// with k = j mod Ls, and writes dst[(j - k) R + k + s Ls].  The tpl threads
// of a line take butterflies jt, jt + tpl, ...
template <typename T, typename T2, int R>
__device__ __forceinline__ void stockham_pass(const T2 *src, T2 *dst, int n, int Ls, float inv_ls,
                                              const T2 *tw, int sign, int jt, int tpl)
{
    const int m = n / R, tstride = n / (Ls * R);
    const FastDiv fl{Ls, inv_ls};
    for (int j = jt; j < m; j += tpl) {
        int k;
        (void)fdivmod(j, fl, k);
        T2 x[R];
#pragma unroll
        for (int q = 0; q < R; q++) x[q] = src[j + q * m];
#pragma unroll
        for (int q = 1; q < R; q++) {
            T2 w = tw[q * k * tstride];
            w.y = sign > 0 ? w.y : -w.y;
            x[q] = cmul(x[q], w);
        }
        dft_small<T, T2, R>(x, sign);
        const int base = (j - k) * R + k;
#pragma unroll
        for (int s = 0; s < R; s++) dst[base + s * Ls] = x[s];
    }
}

// All passes of the plan on the lines of a workgroup (line ll at
// buf[ll * (n + 1)]); each thread works on line lw only.  The result ends in
// the returned buffer.  Every thread of the workgroup must call this.
template <typename T, typename T2>
__device__ T2 *stockham(T2 *buf0, T2 *buf1, const FftPlan &pl, const T2 *tw, int sign, int lw, bool active, int jt,
                        int tpl)
{
    const int n = pl.n, off = lw * (n + 1);
    T2 *src = buf0, *dst = buf1;
    for (int p = 0; p < pl.np; p++) {
        if (active) {
            const int r = pl.rad[p];
            if (r == 4) stockham_pass<T, T2, 4>(src + off, dst + off, n, pl.ls[p], pl.inv_ls[p], tw, sign, jt, tpl);
            else if (r == 2) stockham_pass<T, T2, 2>(src + off, dst + off, n, pl.ls[p], pl.inv_ls[p], tw, sign, jt, tpl);
            else if (r == 3) stockham_pass<T, T2, 3>(src + off, dst + off, n, pl.ls[p], pl.inv_ls[p], tw, sign, jt, tpl);
            else stockham_pass<T, T2, 5>(src + off, dst + off, n, pl.ls[p], pl.inv_ls[p], tw, sign, jt, tpl);
        }
        __syncthreads();
        T2 *t = src;
        src = dst;
        dst = t;
    }
    return src;
}

// LPB strided lines per workgroup: thread t moves elements of line t % LPB,
// so the LPB lines (adjacent x columns) are read and written as contiguous
// runs of LPB elements.
// Loads are issued kLoadBatch at a time into registers before any is stored
// to LDS, so a workgroup waits for one memory round trip per batch rather
// than one per element.
constexpr int kLoadBatch = 8;

template <typename T2, int LPB>
__device__ __forceinline__ void load_lines(const T2 *data, const LineMap &lm, int n, int L0, int nl, T2 *b0)
{
    const int ll = threadIdx.x % LPB;
    if (ll >= nl || !line_ok(lm, L0 + ll)) return;
    const size_t base = line_base<T2>(lm, L0 + ll);
    constexpr int step = kFftThreads / LPB;
    for (int e0 = threadIdx.x / LPB; e0 < n; e0 += kLoadBatch * step) {
        T2 v[kLoadBatch];
#pragma unroll
        for (int u = 0; u < kLoadBatch; u++) {
            const int e = e0 + u * step;
            if (e < n) v[u] = data[base + (size_t)e * lm.se];
        }
#pragma unroll
        for (int u = 0; u < kLoadBatch; u++) {
            const int e = e0 + u * step;
            if (e < n) b0[ll * (n + 1) + e] = v[u];
        }
    }
}

template <typename T2, int LPB>
__device__ __forceinline__ void store_lines(T2 *data, const LineMap &lm, int n, int L0, int nl, const T2 *res)
{
    const int ll = threadIdx.x % LPB;
    if (ll >= nl || !line_ok(lm, L0 + ll)) return;
    const size_t base = line_base<T2>(lm, L0 + ll);
    for (int e = threadIdx.x / LPB; e < n; e += kFftThreads / LPB) data[base + (size_t)e * lm.se] = res[ll * (n + 1) + e];
}

template <typename T2>
__device__ __forceinline__ void load_twiddles(const T2 *tw, int n, T2 *stw)
{
    for (int k = threadIdx.x; k < n; k += kFftThreads) stw[k] = tw[k];
}

// Complex-to-complex transforms of strided lines (the y axis), in place.
template <typename T, int LPB>
__device__ __forceinline__ void c2c_body(const C2CArgs<T> &a, int bid)
{
    using T2 = typename C2T<T>::type;
    extern __shared__ __align__(16) unsigned char smem[];
    const LineMap &lm = a.lm;
    const int n = a.pl.n, ns = n + 1;
    T2 *b0 = (T2 *)smem, *b1 = b0 + LPB * ns, *stw = b1 + LPB * ns;
    const int L0 = bid * LPB, nl = min(LPB, lm.nl - L0);
    load_twiddles(a.tw, n, stw);
    load_lines<T2, LPB>(a.data, lm, n, L0, nl, b0);
    __syncthreads();
    constexpr int TPL = kFftThreads / LPB;
    const int lw = threadIdx.x / TPL;
    T2 *res = stockham<T, T2>(b0, b1, a.pl, stw, a.sign, lw, lw < nl && line_ok(lm, L0 + lw), threadIdx.x % TPL, TPL);
    store_lines<T2, LPB>(a.data, lm, n, L0, nl, res);
}


// The z axis with the influence function in between: forward transform of
// each z line, multiply by theta(m) (ewald_spole.F90:1296-1380) while summing
// the reciprocal stress, inverse transform, all in LDS.  Line L is the column
// (kx, ky) = (a, b) of the half spectrum (see LineMap).
template <typename T, int LPB>
__device__ __forceinline__ void z_conv_body(const ZConvArgs<T> &a, int bid)
{
    using T2 = typename C2T<T>::type;
    extern __shared__ __align__(16) unsigned char smem[];
    double *lred = (double *)smem;  // 6 values x wavefronts
    const LineMap &lm = a.lm;
    const FftPlan &pl = a.pl;
    const InflParams &ip = a.ip;
    const double *__restrict__ n2 = a.n2;
    const int n = pl.n, ns = n + 1;
    T2 *b0 = (T2 *)(smem + kConvRedBytes), *b1 = b0 + LPB * ns, *stw = b1 + LPB * ns;
    const int L0 = bid * LPB, nl = min(LPB, lm.nl - L0);
    load_twiddles(a.tw, n, stw);
    load_lines<T2, LPB>(a.data, lm, n, L0, nl, b0);
    __syncthreads();
    constexpr int TPL = kFftThreads / LPB;
    const int lw = threadIdx.x / TPL, jt = threadIdx.x % TPL;
    const bool active = lw < nl && line_ok(lm, L0 + lw);
    T2 *res = stockham<T, T2>(b0, b1, pl, stw, +1, lw, active, jt, TPL);
    double acc[6] = {0, 0, 0, 0, 0, 0};
    if (active) {
        const double sqrpi = 1.7724538509055159;
        const int L = L0 + lw;
        const int j = L % lm.na_pad, k = L / lm.na_pad;
        const int kk = (2 * k > ip.K[1]) ? k - ip.K[1] : k;
        const double *n2x = n2, *n2y = n2 + ip.K[0], *n2z = n2 + ip.K[0] + ip.K[1];
        const double bxy = n2y[k] * n2x[j];
        const T2 *__restrict__ th = a.theta + line_base<T2>(lm, L);
        const double rcf2 = ip.rcf * ip.rcf;
        // the other half of the spectrum: column j (0 < j < K1/2) stands for
        // its Hermitian partner (-j, -ky, -kz) too, with the same |m| and
        // |Q|^2.  The reference folds an index at K/2 to +K/2 (ewald_spole.F90:
        // 1326-1345), so a partner component at a Nyquist index keeps its sign,
        // which the off-diagonal stress sees.
        const bool partner = !(j == 0 || 2 * j == ip.K[0]);
        const int kkp = (2 * k == ip.K[1]) ? kk : -kk;
        for (int l = jt; l < n; l += TPL) {
            const int ll = (2 * l > ip.K[2]) ? l - ip.K[2] : l;
            const int llp = (2 * l == ip.K[2]) ? ll : -ll;
            const double r3[3] = {ll * ip.rcell[2], ll * ip.rcell[5], ll * ip.rcell[8]};
            const double r2[3] = {r3[0] + kk * ip.rcell[1], r3[1] + kk * ip.rcell[4], r3[2] + kk * ip.rcell[7]};
            const double rp[3] = {r2[0] + j * ip.rcell[0], r2[1] + j * ip.rcell[3], r2[2] + j * ip.rcell[6]};
            const double k2 = rp[0] * rp[0] + rp[1] * rp[1] + rp[2] * rp[2];
            T2 h = res[lw * ns + l];
            if (ip.mode == 1) {
                const double bb1 = n2z[l] * bxy;
                // ewald_general.F90:674-682 with stress_kernel (:1033-1057):
                // B f'(x) pi / x / alpha, f(x) = exp(-x^2)/(sqrt(pi) x^2),
                // f'(x) = -2 f/x - (2/x) exp(-x^2)/sqrt(pi) (spme.F90:421-436)
                const double x = kPi * sqrt(k2) * (1.0 / ip.alpha);
                if (k2 <= ip.cut2 && x > 1.0e-6) {
                    const double ex = exp(-(x * x));
                    const double energy = ex / (sqrpi * (x * x));
                    const double fpd = (-2.0 / x * energy) - ((2.0 / x) * (1.0 / sqrpi) * ex);
                    const double fac = fpd * kPi / x / ip.alpha;
                    h.x = (T)(bb1 * (double)h.x * fac);
                    h.y = (T)(bb1 * (double)h.y * fac);
                } else {
                    h.x = (T)0;
                    h.y = (T)0;
                }
            } else if (const double fac = (double)th[(size_t)l * lm.se].x; fac != 0.0) {
                // theta is zero outside the cutoff sphere and at m = 0;
                // (1 + fp^2)/k2 = 1/k2 + (pi/alpha)^2 with fp = |m| pi/alpha
                const double hr = (double)h.x, hi = (double)h.y;
                const double pv = fac * (hr * hr + hi * hi) * (-2.0 * (1.0 / k2 + rcf2));
                double s0 = rp[0] * rp[0], s1 = rp[0] * rp[1], s2 = rp[0] * rp[2];
                double s3 = rp[1] * rp[1], s4 = rp[1] * rp[2], s5 = rp[2] * rp[2];
                if (partner) {
                    // orthorhombic cells: the partner has the same |m| (and so the same fac)
                    const double q[3] = {-j * ip.rcell[0] + kkp * ip.rcell[1] + llp * ip.rcell[2],
                                         -j * ip.rcell[3] + kkp * ip.rcell[4] + llp * ip.rcell[5],
                                         -j * ip.rcell[6] + kkp * ip.rcell[7] + llp * ip.rcell[8]};
                    s0 += q[0] * q[0];
                    s1 += q[0] * q[1];
                    s2 += q[0] * q[2];
                    s3 += q[1] * q[1];
                    s4 += q[1] * q[2];
                    s5 += q[2] * q[2];
                }
                acc[0] += s0 * pv;
                acc[1] += s1 * pv;
                acc[2] += s2 * pv;
                acc[3] += s3 * pv;
                acc[4] += s4 * pv;
                acc[5] += s5 * pv;
                h.x = (T)(fac * hr);
                h.y = (T)(fac * hi);
            } else {
                h.x = (T)0;
                h.y = (T)0;
            }
            res[lw * ns + l] = h;
        }
    }
    __syncthreads();
    T2 *res2 = stockham<T, T2>(res, res == b0 ? b1 : b0, pl, stw, -1, lw, active, jt, TPL);
    store_lines<T2, LPB>(a.data, lm, n, L0, nl, res2);
    if (ip.mode == 0) block_sum<6>(acc, lred, a.red + (size_t)bid * kReduceSlots + R_RS11);
}

// x axis, forward: real lines 2P and 2P+1 of Q (n contiguous values each)
// packed as one complex line; output half spectra to H (K1h of a row of hs).
// With clear, every grid value is set to zero once read (each is read by one
// thread), so the charge grid is ready for the next spreading.
template <typename T, int LPB>
__device__ __forceinline__ void r2c_body(const R2CArgs<T> &a, int bid)
{
    using T2 = typename C2T<T>::type;
    extern __shared__ __align__(16) unsigned char smem[];
    T *__restrict__ Q = a.Q;
    T2 *__restrict__ H = a.H;
    const int nreal_lines = a.nreal_lines, K1h = a.K1h, hs = a.hs;
    const FftPlan &pl = a.pl;
    const FastDiv div_n = a.div_n, div_h = a.div_h;
    const bool clear = a.clear != 0;
    const int n = pl.n, ns = n + 1;
    T2 *b0 = (T2 *)smem, *b1 = b0 + LPB * ns, *stw = b1 + LPB * ns;
    const int npairs = (nreal_lines + 1) / 2;
    const int P0 = bid * LPB;
    const int np = min(LPB, npairs - P0);
    load_twiddles(a.tw, n, stw);
    for (int i0 = threadIdx.x; i0 < np * n; i0 += kLoadBatch * kFftThreads) {
        T2 z[kLoadBatch];
#pragma unroll
        for (int u = 0; u < kLoadBatch; u++) {
            const int idx = i0 + u * kFftThreads;
            if (idx < np * n) {
                int e;
                const int ll = fdivmod(idx, div_n, e);
                const int a = 2 * (P0 + ll), b = a + 1;
                z[u].x = Q[(size_t)a * n + e];
                z[u].y = b < nreal_lines ? Q[(size_t)b * n + e] : (T)0;
            }
        }
        if (clear) {
#pragma unroll
            for (int u = 0; u < kLoadBatch; u++) {
                const int idx = i0 + u * kFftThreads;
                if (idx < np * n) {
                    int e;
                    const int ll = fdivmod(idx, div_n, e);
                    const int a = 2 * (P0 + ll), b = a + 1;
                    Q[(size_t)a * n + e] = (T)0;
                    if (b < nreal_lines) Q[(size_t)b * n + e] = (T)0;
                }
            }
        }
#pragma unroll
        for (int u = 0; u < kLoadBatch; u++) {
            const int idx = i0 + u * kFftThreads;
            if (idx < np * n) {
                int e;
                const int ll = fdivmod(idx, div_n, e);
                b0[ll * ns + e] = z[u];
            }
        }
    }
    __syncthreads();
    constexpr int TPL = kFftThreads / LPB;
    const int lw = threadIdx.x / TPL;
    T2 *res = stockham<T, T2>(b0, b1, pl, stw, +1, lw, lw < np, threadIdx.x % TPL, TPL);
    for (int idx = threadIdx.x; idx < np * K1h; idx += kFftThreads) {
        int k;
        const int ll = fdivmod(idx, div_h, k);
        const int a = 2 * (P0 + ll), b = a + 1;
        const T2 z = res[ll * ns + k];
        const T2 zc = res[ll * ns + (k == 0 ? 0 : n - k)];
        T2 A, B;
        A.x = (T)0.5 * (z.x + zc.x);
        A.y = (T)0.5 * (z.y - zc.y);
        B.x = (T)0.5 * (z.y + zc.y);
        B.y = (T)-0.5 * (z.x - zc.x);
        H[(size_t)a * hs + k] = A;
        if (b < nreal_lines) H[(size_t)b * hs + k] = B;
    }
}

// x axis, inverse: rebuild the full Hermitian line of two half spectra,
// transform with sign -1 and write the real parts to Q.
template <typename T, int LPB>
__device__ __forceinline__ void c2r_body(const C2RArgs<T> &a, int bid)
{
    using T2 = typename C2T<T>::type;
    extern __shared__ __align__(16) unsigned char smem[];
    const T2 *__restrict__ H = a.H;
    T *__restrict__ Q = a.Q;
    const int nreal_lines = a.nreal_lines, K1h = a.K1h, hs = a.hs;
    const FftPlan &pl = a.pl;
    const FastDiv div_n = a.div_n;
    const int n = pl.n, ns = n + 1;
    T2 *b0 = (T2 *)smem, *b1 = b0 + LPB * ns, *stw = b1 + LPB * ns;
    const int npairs = (nreal_lines + 1) / 2;
    const int P0 = bid * LPB;
    const int np = min(LPB, npairs - P0);
    load_twiddles(a.tw, n, stw);
    for (int i0 = threadIdx.x; i0 < np * n; i0 += kLoadBatch * kFftThreads) {
        T2 A[kLoadBatch], B[kLoadBatch];
#pragma unroll
        for (int u = 0; u < kLoadBatch; u++) {
            const int idx = i0 + u * kFftThreads;
            if (idx < np * n) {
                int k;
                const int ll = fdivmod(idx, div_n, k);
                const int a = 2 * (P0 + ll), b = a + 1;
                const int kh = k < K1h ? k : n - k;
                A[u] = H[(size_t)a * hs + kh];
                if (b < nreal_lines) {
                    B[u] = H[(size_t)b * hs + kh];
                } else {
                    B[u].x = (T)0;
                    B[u].y = (T)0;
                }
            }
        }
#pragma unroll
        for (int u = 0; u < kLoadBatch; u++) {
            const int idx = i0 + u * kFftThreads;
            if (idx < np * n) {
                int k;
                const int ll = fdivmod(idx, div_n, k);
                T2 z;
                if (k < K1h) {
                    z.x = A[u].x - B[u].y;
                    z.y = A[u].y + B[u].x;
                } else {
                    z.x = A[u].x + B[u].y;
                    z.y = -A[u].y + B[u].x;
                }
                b0[ll * ns + k] = z;
            }
        }
    }
    __syncthreads();
    constexpr int TPL = kFftThreads / LPB;
    const int lw = threadIdx.x / TPL;
    T2 *res = stockham<T, T2>(b0, b1, pl, stw, -1, lw, lw < np, threadIdx.x % TPL, TPL);
    for (int idx = threadIdx.x; idx < np * n; idx += kFftThreads) {
        int e;
        const int ll = fdivmod(idx, div_n, e);
        const int a = 2 * (P0 + ll), b = a + 1;
        const T2 z = res[ll * ns + e];
        Q[(size_t)a * n + e] = z.x;
        if (b < nreal_lines) Q[(size_t)b * n + e] = z.y;
    }
}

// The influence function theta(m) = B(m) exp(-pi^2 m^2/alpha^2) /
// (sqrt(pi) (pi |m|/alpha)^2) of ewald_spole.F90:1296-1380 on the half
// spectrum, in the layout of H (x fastest, row stride K1hp), and its stress
// weight theta(m) (-2) (1 + fp^2)/|m|^2 = theta(m) (-2) (1/|m|^2 + (pi/alpha)^2)
// (ewald_spole.F90:1366-1374); both zero outside the spherical cutoff and at
// m = 0.  They depend only on the mesh, the cell and alpha, so the z pass
// reads them instead of forming them per element and step.
template <typename T>
__global__ void __launch_bounds__(256) k_theta(typename C2T<T>::type *__restrict__ th, InflParams ip,
                                               const double *__restrict__ n2, int K1hp)
{
    const long total = (long)ip.K1h * ip.K[1] * ip.K[2];
    const double sqrpi = 1.7724538509055159;
    const double *n2x = n2, *n2y = n2 + ip.K[0], *n2z = n2 + ip.K[0] + ip.K[1];
    for (long e = (long)blockIdx.x * 256 + threadIdx.x; e < total; e += (long)gridDim.x * 256) {
        const int j = (int)(e % ip.K1h);
        const long rest = e / ip.K1h;
        const int k = (int)(rest % ip.K[1]), l = (int)(rest / ip.K[1]);
        const int kk = (2 * k > ip.K[1]) ? k - ip.K[1] : k;
        const int ll = (2 * l > ip.K[2]) ? l - ip.K[2] : l;
        const double r3[3] = {ll * ip.rcell[2], ll * ip.rcell[5], ll * ip.rcell[8]};
        const double r2[3] = {r3[0] + kk * ip.rcell[1], r3[1] + kk * ip.rcell[4], r3[2] + kk * ip.rcell[7]};
        const double rp[3] = {r2[0] + j * ip.rcell[0], r2[1] + j * ip.rcell[3], r2[2] + j * ip.rcell[6]};
        const double bb1 = n2z[l] * (n2y[k] * n2x[j]);
        const double k2 = rp[0] * rp[0] + rp[1] * rp[1] + rp[2] * rp[2];
        double fac = 0.0, wst = 0.0;
        if (k2 <= ip.cut2 && k2 > ip.test_fac) {
            const double fp = sqrt(k2) * ip.rcf;
            fac = bb1 * exp(-(fp * fp)) / (sqrpi * (fp * fp));
            wst = fac * (-2.0 * (1.0 / k2 + ip.rcf * ip.rcf));
        }
        typename C2T<T>::type v;
        v.x = (T)fac;
        v.y = (T)wst;
        th[((size_t)l * ip.K[1] + k) * K1hp + j] = v;
    }
}

int theta_update(Ctx *c, const InflParams &ip, hipStream_t s)
{
    double key[13] = {(double)ip.K[0], (double)ip.K[1], (double)ip.K[2]};
    for (int k = 0; k < 9; k++) key[3 + k] = ip.rcell[k];
    key[12] = ip.alpha;
    if (c->theta_valid && std::equal(key, key + 13, c->theta_key)) return 0;
    const long total = (long)ip.K1h * ip.K[1] * ip.K[2];
    const int nb = (int)std::min<long>(ceil_div(total, 256), 8192);
    if (c->prm.precision == DLPX_MIXED)
        hipLaunchKernelGGL(k_theta<float>, dim3(nb), dim3(256), 0, s, (float2 *)c->theta, ip, c->bsp_n2, c->K1hp);
    else
        hipLaunchKernelGGL(k_theta<double>, dim3(nb), dim3(256), 0, s, (double2 *)c->theta, ip, c->bsp_n2,
                           c->K1hp);
    DLPX_HIP(c, hipGetLastError());
    std::copy(key, key + 13, c->theta_key);
    c->theta_valid = true;
    return 0;
}

// Single-system kernels (one 3D grid) and batched ones (the grids of many
// systems, one argument struct each, see batch_find).  FFT workgroups take
// lines in grid order: no XCD renumbering.
#define DLPX_FFT_KERNELS(NAME, ARGS, BODY)                                                                 \
    template <typename T, int LPB>                                                                         \
    __global__ void __launch_bounds__(kFftThreads) NAME(ARGS<T> a)                                         \
    {                                                                                                      \
        BODY<T, LPB>(a, blockIdx.x);                                                                       \
    }                                                                                                      \
    template <typename T, int LPB>                                                                         \
    __global__ void __launch_bounds__(kFftThreads) NAME##_batch(const ARGS<T> *__restrict__ av,            \
                                                                const int32_t *__restrict__ boff, int nsys) \
    {                                                                                                      \
        const int lb = blockIdx.x;                                                                         \
        if (lb >= boff[nsys]) return;                                                                      \
        const int s = batch_find(boff, nsys, lb);                                                          \
        BODY<T, LPB>(av[s], lb - boff[s]);                                                                 \
    }
DLPX_FFT_KERNELS(k_fft_c2c, C2CArgs, c2c_body)
DLPX_FFT_KERNELS(k_fft_z_conv, ZConvArgs, z_conv_body)
DLPX_FFT_KERNELS(k_fft_x_r2c, R2CArgs, r2c_body)
DLPX_FFT_KERNELS(k_fft_x_c2r, C2RArgs, c2r_body)
#undef DLPX_FFT_KERNELS

// Lines per workgroup: the largest of 16, 8, 4, 2, 1 whose two Stockham
// buffers and twiddles fit 72 KB of LDS (two workgroups per CU); 16 lines give
// 128 B (FP32) or 256 B (FP64) contiguous runs in the y and z passes.
template <typename T>
static size_t fft_lds(int lpb, int n, size_t extra)
{
    return sizeof(typename C2T<T>::type) * ((size_t)2 * lpb * (n + 1) + n) + extra;
}

template <typename T>
static int pick_lpb(int n, size_t extra)
{
    for (int lpb = 16; lpb >= 1; lpb /= 2)
        if (fft_lds<T>(lpb, n, extra) <= 72 * 1024) return lpb;
    return fft_lds<T>(1, n, extra) <= 160 * 1024 ? 1 : 0;
}

#define DLPX_LPB_SWITCH(lpb, CALL) \
    switch (lpb) {                 \
        case 16: CALL(16); break;  \
        case 8: CALL(8); break;    \
        case 4: CALL(4); break;    \
        case 2: CALL(2); break;    \
        default: CALL(1); break;   \
    }

// Lines per workgroup of each pass for one context (0: a length too large
// for LDS).
template <typename T>
static int xy_lpb(int n)
{
    return pick_lpb<T>(n, 0);
}

template <typename T>
static C2CArgs<T> y_args(const Ctx *c, int sign, int lpb)
{
    using T2 = typename C2T<T>::type;
    const int K1h = c->K1h, K2 = c->K[1], K3 = c->K[2];
    C2CArgs<T> a;
    a.lm.na = K1h;
    a.lm.na_pad = (K1h + lpb - 1) / lpb * lpb;
    a.lm.nl = a.lm.na_pad * K3;
    a.lm.sa = 1;
    a.lm.sb = (long)c->K1hp * K2;
    a.lm.se = c->K1hp;
    a.pl = make_plan(K2);
    a.tw = (const T2 *)c->tw + c->K[0];
    a.data = (T2 *)c->H;
    a.sign = sign;
    return a;
}

static LineMap z_lines(const Ctx *c, int lpb)
{
    LineMap lm;
    lm.na = c->K1h;
    lm.na_pad = (c->K1h + lpb - 1) / lpb * lpb;
    lm.nl = lm.na_pad * c->K[1];
    lm.sa = 1;
    lm.sb = c->K1hp;
    lm.se = (long)c->K1hp * c->K[1];
    return lm;
}

template <typename T>
static int z_lpb(const Ctx *c)
{
    return pick_lpb<T>(c->K[2], kConvRedBytes);
}

template <typename T>
static ZConvArgs<T> z_args(const Ctx *c, const InflParams &ip, int lpb)
{
    using T2 = typename C2T<T>::type;
    ZConvArgs<T> a;
    a.lm = z_lines(c, lpb);
    a.pl = make_plan(c->K[2]);
    a.tw = (const T2 *)c->tw + c->K[0] + c->K[1];
    a.data = (T2 *)c->H;
    a.ip = ip;
    a.n2 = c->bsp_n2;
    a.theta = (const typename C2T<T>::type *)c->theta;
    a.red = c->red;
    return a;
}

template <typename T>
static R2CArgs<T> r2c_args(const Ctx *c, void *grid, bool clear)
{
    using T2 = typename C2T<T>::type;
    R2CArgs<T> a;
    const int n = c->K[0];
    a.Q = (T *)grid;
    a.H = (T2 *)c->H;
    a.nreal_lines = c->K[1] * c->K[2];
    a.K1h = c->K1h;
    a.hs = c->K1hp;
    a.pl = make_plan(n);
    a.tw = (const T2 *)c->tw;
    a.div_n = make_div(n);
    a.div_h = make_div(c->K1h);
    a.clear = clear ? 1 : 0;
    return a;
}

template <typename T>
static C2RArgs<T> c2r_args(const Ctx *c, void *grid)
{
    using T2 = typename C2T<T>::type;
    C2RArgs<T> a;
    const int n = c->K[0];
    a.H = (const T2 *)c->H;
    a.Q = (T *)grid;
    a.nreal_lines = c->K[1] * c->K[2];
    a.K1h = c->K1h;
    a.hs = c->K1hp;
    a.pl = make_plan(n);
    a.tw = (const T2 *)c->tw;
    a.div_n = make_div(n);
    return a;
}

static int x_blocks(const Ctx *c, int lpb) { return ceil_div((c->K[1] * c->K[2] + 1) / 2, lpb); }

static bool use_lines(const Ctx *c, int n) { return c->fft_lines && lines_supported(n); }

// the x and y passes fused per z-plane (fft_planes.hip): K1 = K2 of a
// line-kernel length whose plane takes at most 72 KB of LDS (FP32 up to 128,
// FP64 up to 80).  A larger plane leaves a CU no room for a workgroup of the
// pair kernel beside it, and in a resident step the plane kernel then waits
// for the pair kernel's workgroups to drain (measured at 128^3 in FP64: the
// step 0.311 -> 0.319 ms with planes of 132 KB).
template <typename T>
static bool use_plane(const Ctx *c)
{
    return c->fft_planes && use_lines(c, c->K[0]) && c->K[0] == c->K[1] && plane_lds<T>(c->K[0]) <= 72 * 1024;
}

template <typename T>
static PlaneArgs<T> plane_args(const Ctx *c, void *grid, bool clear)
{
    using T2 = typename C2T<T>::type;
    PlaneArgs<T> a;
    a.Q = (T *)grid;
    a.H = (T2 *)c->H;
    a.nplanes = c->K[2];
    a.hs = c->K1hp;
    a.tw = (const T2 *)c->tw;
    a.clear = clear ? 1 : 0;
    return a;
}

template <typename T>
static int run_y(Ctx *c, int sign)
{
    if (use_lines(c, c->K[1])) {
        const int lpb = lines_lpb<T>();
        const C2CArgs<T> a = y_args<T>(c, sign, lpb);
        return lines_c2c<T>(c, c->sp, a, ceil_div(a.lm.nl, lpb));
    }
    const int lpb = xy_lpb<T>(c->K[1]);
    if (!lpb) return fail(c, DLPX_ERR_UNSUPPORTED, "FFT length too large for LDS");
    const C2CArgs<T> a = y_args<T>(c, sign, lpb);
    const size_t lds = fft_lds<T>(lpb, c->K[1], 0);
    const int nb = ceil_div(a.lm.nl, lpb);
#define DLPX_Y(L)                                                                                                   \
    DLPX_LDS_MAX(c, (k_fft_c2c<T, L>));                                                                     \
    hipLaunchKernelGGL((k_fft_c2c<T, L>), dim3(nb), dim3(kFftThreads), lds, c->sp, a)
    DLPX_LPB_SWITCH(lpb, DLPX_Y)
#undef DLPX_Y
    DLPX_HIP(c, hipGetLastError());
    return 0;
}

template <typename T>
static int run_z_conv(Ctx *c, const InflParams &ip)
{
    const int lpb = z_lpb<T>(c);
    if (!lpb) return fail(c, DLPX_ERR_UNSUPPORTED, "FFT length too large for LDS");
    int r;
    if (ip.mode == 0 && (r = theta_update(c, ip, c->sp))) return r;
    if (ip.mode == 0 && use_lines(c, c->K[2])) {
        const int ll = lines_lpb<T>();
        const ZConvArgs<T> a = z_args<T>(c, ip, ll);
        const int nb = ceil_div(a.lm.nl, ll);
        for (int k = R_RS11; k <= R_RS33; k++) c->red_rows[k] = nb;
        return lines_zconv<T>(c, c->sp, a, nb);
    }
    const ZConvArgs<T> a = z_args<T>(c, ip, lpb);
    const size_t lds = fft_lds<T>(lpb, c->K[2], kConvRedBytes);
    const int nb = ceil_div(a.lm.nl, lpb);
#define DLPX_Z(L)                                                                                                   \
    DLPX_LDS_MAX(c, (k_fft_z_conv<T, L>));                                                                     \
    hipLaunchKernelGGL((k_fft_z_conv<T, L>), dim3(nb), dim3(kFftThreads), lds, c->sp, a)
    DLPX_LPB_SWITCH(lpb, DLPX_Z)
#undef DLPX_Z
    if (ip.mode == 0)
        for (int k = R_RS11; k <= R_RS33; k++) c->red_rows[k] = nb;
    DLPX_HIP(c, hipGetLastError());
    return 0;
}

template <typename T>
static int run_x(Ctx *c, bool forward, void *grid, bool clear = false)
{
    const int n = c->K[0];
    if (use_lines(c, n)) {
        const int ll = lines_lpb<T>(), nb = x_blocks(c, ll);
        if (forward) {
            const R2CArgs<T> a = r2c_args<T>(c, grid, clear);
            return lines_x<T>(c, c->sp, &a, nullptr, nb);
        }
        const C2RArgs<T> a = c2r_args<T>(c, grid);
        return lines_x<T>(c, c->sp, nullptr, &a, nb);
    }
    const int lpb = xy_lpb<T>(n);
    if (!lpb) return fail(c, DLPX_ERR_UNSUPPORTED, "FFT length too large for LDS");
    const size_t lds = fft_lds<T>(lpb, n, 0);
    const int nb = x_blocks(c, lpb);
    if (forward) {
        const R2CArgs<T> a = r2c_args<T>(c, grid, clear);
#define DLPX_R2C(L)                                                                                                 \
    DLPX_LDS_MAX(c, (k_fft_x_r2c<T, L>));                                                                     \
    hipLaunchKernelGGL((k_fft_x_r2c<T, L>), dim3(nb), dim3(kFftThreads), lds, c->sp, a)
        DLPX_LPB_SWITCH(lpb, DLPX_R2C)
#undef DLPX_R2C
    } else {
        const C2RArgs<T> a = c2r_args<T>(c, grid);
#define DLPX_C2R(L)                                                                                                 \
    DLPX_LDS_MAX(c, (k_fft_x_c2r<T, L>));                                                                     \
    hipLaunchKernelGGL((k_fft_x_c2r<T, L>), dim3(nb), dim3(kFftThreads), lds, c->sp, a)
        DLPX_LPB_SWITCH(lpb, DLPX_C2R)
#undef DLPX_C2R
    }
    DLPX_HIP(c, hipGetLastError());
    return 0;
}

template <typename T>
static int convolve_t(Ctx *c, const InflParams &ip, void *in, void *out, bool clear_in)
{
    int r;
    if (use_plane<T>(c)) {
        if ((r = plane_run<T>(c, c->sp, false, c->K[0], plane_args<T>(c, in, clear_in)))) return r;
        if ((r = run_z_conv<T>(c, ip))) return r;
        return plane_run<T>(c, c->sp, true, c->K[0], plane_args<T>(c, out, false));
    }
    if ((r = run_x<T>(c, true, in, clear_in))) return r;
    if ((r = run_y<T>(c, +1))) return r;
    if ((r = run_z_conv<T>(c, ip))) return r;
    if ((r = run_y<T>(c, -1))) return r;
    return run_x<T>(c, false, out);
}

int fft_convolve(Ctx *c, const InflParams &ip, void *in, void *out, bool clear_in)
{
    return c->prm.precision == DLPX_MIXED ? convolve_t<float>(c, ip, in, out, clear_in)
                                          : convolve_t<double>(c, ip, in, out, clear_in)
;
}

constexpr int kPlaneKey = 1 << 20;                 // group keys of the plane kernels: kPlaneKey + K1
constexpr int kGroupPlaneFwd = 5, kGroupPlaneInv = 6;  // FftGroup::kind of the plane kernels

// Batched convolution Q -> Phi (influence function mode 0, the forward pass
// clearing Q) over the systems with electrostatics: each pass is one launch
// per group of systems, a group for each line length that has compiled line
// kernels and one for the rest (generic kernels, with one lines-per-workgroup
// value per pass: the smallest any of its systems needs, so every system's
// lines fit its LDS).
template <typename T, typename A, typename Build>
static int add_group(const std::vector<Ctx *> &cs, FftBatch &fb, int kind, int key, int lpb,
                     const std::vector<int> &keys, Build build, int gkind = -1, int gn = -1)
{
    const size_t ns = cs.size();
    std::vector<A> args(ns, A{});
    std::vector<int> nb(ns, 0);
    size_t lds = 0;
    for (size_t k = 0; k < ns; k++) {
        if (keys[k] != key) continue;
        size_t l = 0;
        args[k] = build(cs[k], k, lpb, nb[k], l);
        lds = std::max(lds, l);
        if (kind == 2) fb.conv_rows[k] = nb[k];
    }
    FftGroup g;
    g.kind = gkind >= 0 ? gkind : kind;
    g.n = gn >= 0 ? gn : (key > 0 ? key : 0);
    g.lpb = lpb;
    int r;
    if ((r = batch_upload(cs[0], g.bl, args.data(), sizeof(A), nb, false))) return r;
    g.bl.lds = lds;
    fb.groups.push_back(g);
    return 0;
}

template <typename T>
static int fft_prepare_t(const std::vector<Ctx *> &cs, const std::vector<InflParams> &ips, FftBatch &fb)
{
    const int cap = fb.lpb_cap > 0 ? fb.lpb_cap : 16;
    for (auto &g : fb.groups) batch_free(g.bl);
    fb.groups.clear();
    const size_t ns = cs.size();
    fb.conv_rows.assign(ns, 0);
    int lg[3] = {cap, cap, cap};  // generic lines per workgroup per axis
    for (size_t k = 0; k < ns; k++) {
        Ctx *c = cs[k];
        if (!c->prm.electrostatics) continue;
        int r;
        if ((r = theta_update(c, ips[k], c->stream))) return r;
        DLPX_HIP(c, hipStreamSynchronize(c->stream));
        for (int d = 0; d < 3; d++) {
            if (use_lines(c, c->K[d])) continue;
            const int l = d == 2 ? z_lpb<T>(c) : xy_lpb<T>(c->K[d]);
            if (!l) return fail(c, DLPX_ERR_UNSUPPORTED, "FFT length too large for LDS");
            lg[d] = std::min(lg[d], l);
        }
    }
    const int axis_of[5] = {0, 1, 2, 1, 0};
    for (int kind = 0; kind < 5; kind++) {
        const int d = axis_of[kind];
        // group key per system: its line length with line kernels, 0 for the
        // generic group, -1 without a mesh
        // (kPlaneKey + K1 for the fused x-y plane kernels, which also take the
        // y passes)
        std::vector<int> keys(ns, -1), distinct;
        for (size_t k = 0; k < ns; k++) {
            Ctx *c = cs[k];
            if (!c->prm.electrostatics) continue;
            if (use_plane<T>(c)) {
                if (kind == 1 || kind == 3) continue;
                if (kind != 2) {
                    keys[k] = kPlaneKey + c->K[0];
                    if (std::find(distinct.begin(), distinct.end(), keys[k]) == distinct.end()) distinct.push_back(keys[k]);
                    continue;
                }
            }
            keys[k] = use_lines(c, c->K[d]) ? c->K[d] : 0;
            if (std::find(distinct.begin(), distinct.end(), keys[k]) == distinct.end()) distinct.push_back(keys[k]);
        }
        for (int key : distinct) {
            const int lpb = key > 0 ? lines_lpb<T>() : lg[d];
            int r = 0;
            if (key > kPlaneKey) {
                const int n = key - kPlaneKey;
                r = add_group<T, PlaneArgs<T>>(
                    cs, fb, kind, key, 1, keys,
                    [&](Ctx *c, size_t, int, int &nb, size_t &lds) {
                        nb = c->K[2];
                        lds = plane_lds<T>(n);
                        return kind == 0 ? plane_args<T>(c, c->Q, true) : plane_args<T>(c, c->Phi, false);
                    },
                    kind == 0 ? kGroupPlaneFwd : kGroupPlaneInv, n);
            } else if (kind == 0)
                r = add_group<T, R2CArgs<T>>(cs, fb, kind, key, lpb, keys,
                                             [&](Ctx *c, size_t, int l, int &nb, size_t &lds) {
                                                 nb = x_blocks(c, l);
                                                 lds = key > 0 ? lines_lds<T>(2, key) : fft_lds<T>(l, c->K[0], 0);
                                                 return r2c_args<T>(c, c->Q, true);
                                             });
            else if (kind == 4)
                r = add_group<T, C2RArgs<T>>(cs, fb, kind, key, lpb, keys,
                                             [&](Ctx *c, size_t, int l, int &nb, size_t &lds) {
                                                 nb = x_blocks(c, l);
                                                 lds = key > 0 ? lines_lds<T>(3, key) : fft_lds<T>(l, c->K[0], 0);
                                                 return c2r_args<T>(c, c->Phi);
                                             });
            else if (kind == 2)
                r = add_group<T, ZConvArgs<T>>(cs, fb, kind, key, lpb, keys,
                                               [&](Ctx *c, size_t k, int l, int &nb, size_t &lds) {
                                                   const ZConvArgs<T> a = z_args<T>(c, ips[k], l);
                                                   nb = ceil_div(a.lm.nl, l);
                                                   lds = key > 0 ? lines_lds<T>(1, key)
                                                                 : fft_lds<T>(l, c->K[2], kConvRedBytes);
                                                   return a;
                                               });
            else
                r = add_group<T, C2CArgs<T>>(cs, fb, kind, key, lpb, keys,
                                             [&](Ctx *c, size_t, int l, int &nb, size_t &lds) {
                                                 const C2CArgs<T> a = y_args<T>(c, kind == 1 ? +1 : -1, l);
                                                 nb = ceil_div(a.lm.nl, l);
                                                 lds = key > 0 ? lines_lds<T>(0, key) : fft_lds<T>(l, c->K[1], 0);
                                                 return a;
                                             });
            if (r) return r;
        }
    }
    return 0;
}

int fft_prepare(const std::vector<Ctx *> &cs, const std::vector<InflParams> &ips, FftBatch &fb)
{
    return cs[0]->prm.precision == DLPX_MIXED ? fft_prepare_t<float>(cs, ips, fb) : fft_prepare_t<double>(cs, ips, fb);
}

template <typename T>
static int fft_launch_t(Ctx *c, hipStream_t s, const FftBatch &fb)
{
    for (const FftGroup &g : fb.groups) {
        if (g.bl.total == 0) continue;
        if (g.kind == kGroupPlaneFwd || g.kind == kGroupPlaneInv) {
            int r;
            if ((r = plane_batch<T>(c, s, g.kind == kGroupPlaneInv, g.n, g.bl))) return r;
            continue;
        }
        if (g.n > 0) {
            int r;
            if ((r = lines_batch<T>(c, s, g.kind, g.n, g.bl))) return r;
            continue;
        }
        const BatchLaunch &bl = g.bl;
#define DLPX_FB(KERNEL, ARGS, L)                                                                                 \
    DLPX_LDS_MAX(c, (KERNEL<T, L>));                                                                             \
    hipLaunchKernelGGL((KERNEL<T, L>), dim3(bl.total), dim3(kFftThreads), bl.lds, s, (const ARGS<T> *)bl.args,   \
                       bl.boff, bl.nsys)
#define DLPX_FB_R2C(L) DLPX_FB(k_fft_x_r2c_batch, R2CArgs, L)
#define DLPX_FB_Y(L) DLPX_FB(k_fft_c2c_batch, C2CArgs, L)
#define DLPX_FB_Z(L) DLPX_FB(k_fft_z_conv_batch, ZConvArgs, L)
#define DLPX_FB_C2R(L) DLPX_FB(k_fft_x_c2r_batch, C2RArgs, L)
        switch (g.kind) {
            case 0: DLPX_LPB_SWITCH(g.lpb, DLPX_FB_R2C) break;
            case 2: DLPX_LPB_SWITCH(g.lpb, DLPX_FB_Z) break;
            case 4: DLPX_LPB_SWITCH(g.lpb, DLPX_FB_C2R) break;
            default: DLPX_LPB_SWITCH(g.lpb, DLPX_FB_Y) break;
        }
#undef DLPX_FB_R2C
#undef DLPX_FB_Y
#undef DLPX_FB_Z
#undef DLPX_FB_C2R
#undef DLPX_FB
    }
    DLPX_HIP(c, hipGetLastError());
    return 0;
}

int fft_launch(Ctx *c, hipStream_t s, const FftBatch &fb)
{
    return c->prm.precision == DLPX_MIXED ? fft_launch_t<float>(c, s, fb) : fft_launch_t<double>(c, s, fb);
}

int conv_blocks(const Ctx *c)
{
    if (use_lines(c, c->K[2])) {
        const int ll = c->prm.precision == DLPX_MIXED ? lines_lpb<float>() : lines_lpb<double>();
        return ceil_div((long)z_lines(c, ll).nl, ll);
    }
    const int lpb = c->prm.precision == DLPX_MIXED ? z_lpb<float>(c) : z_lpb<double>(c);
    return lpb ? ceil_div((long)z_lines(c, lpb).nl, lpb) : 0;
}

}  // namespace dlpx
