// The x and y passes of the SPME convolution fused per z-plane, for meshes
// with K1 = K2 = N of a line-kernel length.  One workgroup holds one plane of
// the half spectrum (N rows of N/2 + 1 complex values) in LDS:
//
//   forward: the plane's N real rows are read once (row pairs packed as
//            complex lines), transformed along x, split into half spectra
//            into the LDS plane, transformed along y per column, and the
//            result written to the half-spectrum grid;
//   inverse: the half-spectrum plane is read once, transformed along y per
//            column in LDS, recombined per row pair into Hermitian lines,
//            transformed along x and written back as real rows.
//
// So the x and y passes cost one read and one write of the grid instead of
// two of each, and one launch instead of two.  The line transforms are the
// two-stage register DFTs of the line kernels (fft_reg.h); every exchange is
// made in place in the plane buffer, between barriers.  Same convention as
// gpfa235.F90:72 (sign +1 forward, no normalisation).
#include <algorithm>

#include "fft_reg.h"

namespace dlpx {

// Threads of a plane workgroup: one per (column, thread-of-line) pair of the
// y pass, which covers the (row pair, thread-of-line) pairs of the x pass.
template <int N, int N2>
constexpr int plane_threads()
{
    return ((N / 2 + 1) * N2 + 63) / 64 * 64;
}

// Complex elements of the plane buffer: the x pass's row pairs with a
// padded line stride, or the half-spectrum plane, whichever is larger.
constexpr int plane_elems(int n)
{
    return std::max(((n + 1) / 2) * (n + 1), n * (n / 2 + 1));
}

template <typename T, int N1, int N2>
__device__ __forceinline__ void plane_fwd_body(const PlaneArgs<T> &a, int z)
{
    using T2 = typename C2T<T>::type;
    constexpr int N = N1 * N2, EXS = N + 1, RB = (N1 + N2 - 1) / N2, K1h = N / 2 + 1, NP = (N + 1) / 2;
    constexpr int NT = plane_threads<N, N2>();
    constexpr int UN = (NP * N + NT - 1) / NT, UH = (NP * K1h + NT - 1) / NT;
    extern __shared__ __align__(16) unsigned char smem[];
    T2 *buf = (T2 *)smem, *stw = buf + plane_elems(N);
    const int tid = threadIdx.x;
    T *__restrict__ Qp = a.Q + (size_t)z * N * N;
    // 1. rows 2l and 2l+1 as the real and imaginary parts of complex line l,
    //    read in whole rows; the loads are unconditional (clamped indices,
    //    values of absent slots dropped) so that all are in flight together
    {
        T2 v[UN];
#pragma unroll
        for (int u = 0; u < UN; u++) {
            const int idx = tid + u * NT;
            const int l2 = min(idx / N, NP - 1), e = idx % N;
            v[u].x = Qp[(2 * l2) * N + e];
            v[u].y = Qp[min(2 * l2 + 1, N - 1) * N + e];
        }
#pragma unroll
        for (int u = 0; u < UN; u++) {
            const int idx = tid + u * NT;
            const int l2 = idx / N, e = idx % N;
            if (l2 < NP) {
                T2 w = v[u];
                if (2 * l2 + 1 >= N) w.y = (T)0;
                buf[l2 * EXS + e] = w;
                if (a.clear) {
                    Qp[(2 * l2) * N + e] = (T)0;
                    if (2 * l2 + 1 < N) Qp[(2 * l2 + 1) * N + e] = (T)0;
                }
            }
        }
    }
    for (int k = tid; k < N; k += NT) stw[k] = a.tw[k];
    __syncthreads();
    // 2. x transforms: line ll on threads t < N2
    {
        const int ll = tid % NP, t = tid / NP;
        const bool on = t < N2;
        T2 x[N1];
        if (on)
#pragma unroll
            for (int n1 = 0; n1 < N1; n1++) x[n1] = buf[ll * EXS + t + N2 * n1];
        __syncthreads();
        if (on) stage_a<T, N1, N2, 1>(x, t, buf + ll * EXS, stw);
        __syncthreads();
        T2 y[RB][N2];
        if (on)
#pragma unroll
            for (int r = 0; r < RB; r++)
                if (t + N2 * r < N1) stage_b<T, N1, N2, 1>(t + N2 * r, buf + ll * EXS, y[r]);
        __syncthreads();
        if (on)
#pragma unroll
            for (int r = 0; r < RB; r++) {
                const int m1 = t + N2 * r;
                if (m1 < N1)
#pragma unroll
                    for (int p = 0; p < N2; p++) buf[ll * EXS + m1 + N1 * RegDft<T, T2, N2, 1>::perm(p)] = y[r][p];
            }
        __syncthreads();
    }
    // 3. the two half spectra of each line: row 2l gets (Z_k + conj Z_{N-k})/2,
    //    row 2l+1 (Z_k - conj Z_{N-k})/(2i); the plane layout [row][k]
    //    overwrites the lines, so everything is read first
    {
        T2 hv[UH][2];
#pragma unroll
        for (int u = 0; u < UH; u++) {
            const int idx = tid + u * NT;
            const int l2 = min(idx / K1h, NP - 1), k = idx % K1h;
            const T2 zz = buf[l2 * EXS + k], zc = buf[l2 * EXS + (k == 0 ? 0 : N - k)];
            hv[u][0].x = (T)0.5 * (zz.x + zc.x);
            hv[u][0].y = (T)0.5 * (zz.y - zc.y);
            hv[u][1].x = (T)0.5 * (zz.y + zc.y);
            hv[u][1].y = (T)-0.5 * (zz.x - zc.x);
        }
        __syncthreads();
#pragma unroll
        for (int u = 0; u < UH; u++) {
            const int idx = tid + u * NT;
            const int l2 = idx / K1h, k = idx % K1h;
            if (l2 < NP) {
                buf[(2 * l2) * K1h + k] = hv[u][0];
                if (2 * l2 + 1 < N) buf[(2 * l2 + 1) * K1h + k] = hv[u][1];
            }
        }
        __syncthreads();
    }
    // 4. y transforms: column c on threads t < N2, exchanged in place in the
    //    column (element stride K1h), the result to the half-spectrum grid
    {
        const int c = tid % K1h, t = tid / K1h;
        const bool on = t < N2;
        T2 *col = buf + c;
        T2 x[N1];
        if (on)
#pragma unroll
            for (int n1 = 0; n1 < N1; n1++) x[n1] = col[(t + N2 * n1) * K1h];
        __syncthreads();
        if (on) stage_a<T, N1, N2, 1, K1h>(x, t, col, stw);
        __syncthreads();
        if (on) {
            T2 *__restrict__ Hp = a.H + (size_t)z * N * a.hs + c;
#pragma unroll
            for (int r = 0; r < RB; r++) {
                const int m1 = t + N2 * r;
                if (m1 < N1) {
                    T2 y[N2];
                    stage_b<T, N1, N2, 1, K1h>(m1, col, y);
#pragma unroll
                    for (int p = 0; p < N2; p++)
                        Hp[(size_t)(m1 + N1 * RegDft<T, T2, N2, 1>::perm(p)) * a.hs] = y[p];
                }
            }
        }
    }
}

template <typename T, int N1, int N2>
__device__ __forceinline__ void plane_inv_body(const PlaneArgs<T> &a, int z)
{
    using T2 = typename C2T<T>::type;
    constexpr int N = N1 * N2, EXS = N + 1, RB = (N1 + N2 - 1) / N2, K1h = N / 2 + 1, NP = (N + 1) / 2;
    constexpr int NT = plane_threads<N, N2>();
    constexpr int UL = (N * K1h + NT - 1) / NT, UN = (NP * N + NT - 1) / NT;
    extern __shared__ __align__(16) unsigned char smem[];
    T2 *buf = (T2 *)smem, *stw = buf + plane_elems(N);
    const int tid = threadIdx.x;
    // 1. the half-spectrum plane into LDS, layout [row][k] (unconditional
    //    clamped loads)
    {
        const T2 *__restrict__ Hp = a.H + (size_t)z * N * a.hs;
        T2 v[UL];
#pragma unroll
        for (int u = 0; u < UL; u++) {
            const int idx = min(tid + u * NT, N * K1h - 1);
            const int row = idx / K1h, k = idx % K1h;
            v[u] = Hp[(size_t)row * a.hs + k];
        }
#pragma unroll
        for (int u = 0; u < UL; u++) {
            const int idx = tid + u * NT;
            if (idx < N * K1h) buf[idx] = v[u];
        }
    }
    for (int k = tid; k < N; k += NT) stw[k] = a.tw[k];
    __syncthreads();
    // 2. inverse y transforms per column, in place, natural order back
    {
        const int c = tid % K1h, t = tid / K1h;
        const bool on = t < N2;
        T2 *col = buf + c;
        T2 x[N1];
        if (on)
#pragma unroll
            for (int n1 = 0; n1 < N1; n1++) x[n1] = col[(t + N2 * n1) * K1h];
        __syncthreads();
        if (on) stage_a<T, N1, N2, -1, K1h>(x, t, col, stw);
        __syncthreads();
        T2 y[RB][N2];
        if (on)
#pragma unroll
            for (int r = 0; r < RB; r++)
                if (t + N2 * r < N1) stage_b<T, N1, N2, -1, K1h>(t + N2 * r, col, y[r]);
        __syncthreads();
        if (on)
#pragma unroll
            for (int r = 0; r < RB; r++) {
                const int m1 = t + N2 * r;
                if (m1 < N1)
#pragma unroll
                    for (int p = 0; p < N2; p++) col[(m1 + N1 * RegDft<T, T2, N2, -1>::perm(p)) * K1h] = y[r][p];
            }
        __syncthreads();
    }
    // 3. rows 2l and 2l+1 (half spectra A and B) as the Hermitian line
    //    A + iB, transformed along x; its real and imaginary parts are the two
    //    real rows
    {
        const int ll = tid % NP, t = tid / NP;
        const bool on = t < N2;
        T2 x[N1];
        if (on) {
            const T2 *ra = buf + (2 * ll) * K1h, *rb = buf + min(2 * ll + 1, N - 1) * K1h;
            const bool hb = 2 * ll + 1 < N;
#pragma unroll
            for (int n1 = 0; n1 < N1; n1++) {
                const int k = t + N2 * n1;
                const int kh = k < K1h ? k : N - k;
                const T2 A = ra[kh];
                T2 B = rb[kh];
                if (!hb) B = zero2<T2>();
                T2 w;
                if (k < K1h) {
                    w.x = A.x - B.y;
                    w.y = A.y + B.x;
                } else {
                    w.x = A.x + B.y;
                    w.y = -A.y + B.x;
                }
                x[n1] = w;
            }
        }
        __syncthreads();
        if (on) stage_a<T, N1, N2, -1>(x, t, buf + ll * EXS, stw);
        __syncthreads();
        T2 y[RB][N2];
        if (on)
#pragma unroll
            for (int r = 0; r < RB; r++)
                if (t + N2 * r < N1) stage_b<T, N1, N2, -1>(t + N2 * r, buf + ll * EXS, y[r]);
        __syncthreads();
        if (on)
#pragma unroll
            for (int r = 0; r < RB; r++) {
                const int m1 = t + N2 * r;
                if (m1 < N1)
#pragma unroll
                    for (int p = 0; p < N2; p++) buf[ll * EXS + m1 + N1 * RegDft<T, T2, N2, -1>::perm(p)] = y[r][p];
            }
        __syncthreads();
    }
    // 4. the real rows, written in whole rows
    T *__restrict__ Qp = a.Q + (size_t)z * N * N;
#pragma unroll
    for (int u = 0; u < UN; u++) {
        const int idx = tid + u * NT;
        const int l2 = idx / N, e = idx % N;
        if (l2 < NP) {
            const T2 w = buf[l2 * EXS + e];
            Qp[(2 * l2) * N + e] = w.x;
            if (2 * l2 + 1 < N) Qp[(2 * l2 + 1) * N + e] = w.y;
        }
    }
}

template <typename T, int N1, int N2, bool INV>
__global__ void __launch_bounds__(((N1 * N2 / 2 + 1) * N2 + 63) / 64 * 64) k_plane(PlaneArgs<T> a)
{
    if ((int)blockIdx.x >= a.nplanes) return;
    if (INV) plane_inv_body<T, N1, N2>(a, blockIdx.x);
    else plane_fwd_body<T, N1, N2>(a, blockIdx.x);
}
template <typename T, int N1, int N2, bool INV>
__global__ void __launch_bounds__(((N1 * N2 / 2 + 1) * N2 + 63) / 64 * 64)
    k_plane_batch(const PlaneArgs<T> *__restrict__ av, const int32_t *__restrict__ boff, int nsys)
{
    const int lb = blockIdx.x;
    if (lb >= boff[nsys]) return;
    const int s = batch_find(boff, nsys, lb);
    if (INV) plane_inv_body<T, N1, N2>(av[s], lb - boff[s]);
    else plane_fwd_body<T, N1, N2>(av[s], lb - boff[s]);
}

// ---- host side -------------------------------------------------------------
template <typename T>
size_t plane_lds(int n)
{
    return sizeof(typename C2T<T>::type) * (size_t)(plane_elems(n) + n);
}
template size_t plane_lds<float>(int);
template size_t plane_lds<double>(int);

template <typename T>
int plane_run(Ctx *c, hipStream_t s, bool inverse, int n, const PlaneArgs<T> &a)
{
    const size_t lds = plane_lds<T>(n);
    switch (n) {
#define DLPX_CASE(N, A, B)                                                                                        \
    case N:                                                                                                       \
        if (inverse) {                                                                                            \
            DLPX_LDS_MAX(c, (k_plane<T, A, B, true>));                                                            \
            hipLaunchKernelGGL((k_plane<T, A, B, true>), dim3(a.nplanes), dim3(plane_threads<N, B>()), lds, s, a); \
        } else {                                                                                                  \
            DLPX_LDS_MAX(c, (k_plane<T, A, B, false>));                                                           \
            hipLaunchKernelGGL((k_plane<T, A, B, false>), dim3(a.nplanes), dim3(plane_threads<N, B>()), lds, s,   \
                               a);                                                                                \
        }                                                                                                         \
        break;
        DLPX_LINE_LENGTHS(DLPX_CASE)
#undef DLPX_CASE
        default: return fail(c, DLPX_ERR_STATE, "no plane kernel for this FFT length");
    }
    DLPX_HIP(c, hipGetLastError());
    return 0;
}

template <typename T>
int plane_batch(Ctx *c, hipStream_t s, bool inverse, int n, const BatchLaunch &bl)
{
    const dim3 g(bl.total);
    switch (n) {
#define DLPX_CASE(N, A, B)                                                                                        \
    case N:                                                                                                       \
        if (inverse) {                                                                                            \
            DLPX_LDS_MAX(c, (k_plane_batch<T, A, B, true>));                                                      \
            hipLaunchKernelGGL((k_plane_batch<T, A, B, true>), g, dim3(plane_threads<N, B>()), bl.lds, s,         \
                               (const PlaneArgs<T> *)bl.args, bl.boff, bl.nsys);                                  \
        } else {                                                                                                  \
            DLPX_LDS_MAX(c, (k_plane_batch<T, A, B, false>));                                                     \
            hipLaunchKernelGGL((k_plane_batch<T, A, B, false>), g, dim3(plane_threads<N, B>()), bl.lds, s,        \
                               (const PlaneArgs<T> *)bl.args, bl.boff, bl.nsys);                                  \
        }                                                                                                         \
        break;
        DLPX_LINE_LENGTHS(DLPX_CASE)
#undef DLPX_CASE
        default: return fail(c, DLPX_ERR_STATE, "no plane kernel for this FFT length");
    }
    DLPX_HIP(c, hipGetLastError());
    return 0;
}

template int plane_run<float>(Ctx *, hipStream_t, bool, int, const PlaneArgs<float> &);
template int plane_run<double>(Ctx *, hipStream_t, bool, int, const PlaneArgs<double> &);
template int plane_batch<float>(Ctx *, hipStream_t, bool, int, const BatchLaunch &);
template int plane_batch<double>(Ctx *, hipStream_t, bool, int, const BatchLaunch &);

}  // namespace dlpx
