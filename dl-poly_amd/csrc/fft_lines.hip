// Line transforms of compile-time length N = N1 x N2 for the SPME
// convolution (the lengths of the meshes in common use; any other length
// runs the generic Stockham kernels of fft.hip).  Same convention as
// gpfa235.F90:72: sign +1 is sum_k c_k exp(+2 pi i j k/n), no normalisation.
//
// A line is split as element n = t + N2 n1 (t < N2, n1 < N1): thread t of
// the line loads its N1 elements straight from the grid into registers,
//   stage A: DFT over n1 in registers, twiddle by exp(sign 2 pi i t m1 / N),
//            one LDS exchange,
//   stage B: DFT over t in registers for each m1 the thread owns, giving
//            X[m1 + N1 m2],
// so a line costs one LDS round trip and one barrier instead of a Stockham
// pass per radix.  The in-register DFTs are built at compile time from
// radix 2, 3, 4 and 5 butterflies with constant twiddles.  The z pass runs
// forward, influence function and inverse (the inverse in the transposed
// order, starting from the registers stage B left) with two exchanges.
#include <algorithm>

#include "fft_reg.h"

namespace dlpx {

// Lane mapping of the y and z passes: lane ll = threadIdx % LPB is the line
// (adjacent x columns, so each element row is one 128-byte run), t =
// threadIdx / LPB its thread.  Padding lanes (t >= N2) and lanes of absent
// lines read line 0's elements through a clamped address, so every load is
// unconditional and all of a thread's loads are in flight together; their
// values are discarded.
template <typename T, int N1, int N2>
__device__ __forceinline__ void load_line(const typename C2T<T>::type *__restrict__ data, size_t base, long se,
                                          int t, bool on, typename C2T<T>::type (&x)[N1])
{
    using T2 = typename C2T<T>::type;
    const int tc = min(t, N2 - 1);
#pragma unroll
    for (int n1 = 0; n1 < N1; n1++) x[n1] = data[base + (size_t)(tc + N2 * n1) * se];
    if (!on)
#pragma unroll
        for (int n1 = 0; n1 < N1; n1++) x[n1] = zero2<T2>();
}

// ---- y axis: complex lines of the half spectrum, in place --------------------
template <typename T, int N1, int N2, int LPB, int SGN>
__device__ __forceinline__ void lines_c2c_body(const C2CArgs<T> &a, int bid)
{
    using T2 = typename C2T<T>::type;
    constexpr int N = N1 * N2, EXS = N + 1, RB = (N1 + N2 - 1) / N2;
    extern __shared__ __align__(16) unsigned char smem[];
    T2 *ex = (T2 *)smem, *stw = ex + LPB * EXS;
    const LineMap &lm = a.lm;
    const int ll = threadIdx.x % LPB, t = threadIdx.x / LPB;
    const int L = bid * LPB + ll;
    const bool on = t < N2 && line_ok(lm, L);
    const size_t base = on ? line_base<T2>(lm, L) : 0;
    T2 x[N1];
    load_line<T, N1, N2>(a.data, base, lm.se, t, on, x);
    for (int k = threadIdx.x; k < N; k += blockDim.x) stw[k] = a.tw[k];
    __syncthreads();
    T2 *exl = ex + ll * EXS;
    if (t < N2) stage_a<T, N1, N2, SGN>(x, t, exl, stw);
    __syncthreads();
    if (!on) return;
#pragma unroll
    for (int r = 0; r < RB; r++) {
        const int m1 = t + N2 * r;
        if (m1 < N1) {
            T2 y[N2];
            stage_b<T, N1, N2, SGN>(m1, exl, y);
#pragma unroll
            for (int p = 0; p < N2; p++)
                a.data[base + (size_t)(m1 + N1 * RegDft<T, T2, N2, SGN>::perm(p)) * lm.se] = y[p];
        }
    }
}

// ---- z axis with the influence function (mode 0) ----------------------------
template <typename T, int N1, int N2, int LPB>
__device__ __forceinline__ void lines_zconv_body(const ZConvArgs<T> &a, int bid)
{
    using T2 = typename C2T<T>::type;
    constexpr int N = N1 * N2, EXS = N + 1, RB = (N1 + N2 - 1) / N2;
    using D2 = RegDft<T, T2, N2, 1>;
    extern __shared__ __align__(16) unsigned char smem[];
    double *lred = (double *)smem;
    T2 *ex0 = (T2 *)(smem + kConvRedBytes), *ex1 = ex0 + LPB * EXS, *stw = ex1 + LPB * EXS;
    const LineMap &lm = a.lm;
    const InflParams &ip = a.ip;
    const int ll = threadIdx.x % LPB, t = threadIdx.x / LPB;
    const int L = bid * LPB + ll;
    const bool on = t < N2 && line_ok(lm, L);
    const size_t base = on ? line_base<T2>(lm, L) : 0;
    T2 x[N1];
    load_line<T, N1, N2>(a.data, base, lm.se, t, on, x);
    // theta and its stress weight for every element this thread will hold
    // after the forward transform (X[m1 + N1 m2], m1 = t + N2 r), fetched
    // with the line
    T2 thv[RB][N2];
    {
        const T2 *__restrict__ th = a.theta + base;
        const int tc = min(t, N2 - 1);
#pragma unroll
        for (int r = 0; r < RB; r++)
#pragma unroll
            for (int p = 0; p < N2; p++) {
                const int m1 = min(tc + N2 * r, N1 - 1);
                thv[r][p] = th[(size_t)(m1 + N1 * D2::perm(p)) * lm.se];
            }
    }
    for (int k = threadIdx.x; k < N; k += blockDim.x) stw[k] = a.tw[k];
    __syncthreads();
    if (t < N2) stage_a<T, N1, N2, 1>(x, t, ex0 + ll * EXS, stw);
    __syncthreads();
    double acc[6] = {0, 0, 0, 0, 0, 0};
    if (t < N2) {
        // Reciprocal stress (ewald_spole.F90:1362-1378) of an orthorhombic
        // cell: m = (j r0, ky r4, kz r8) along the line, so with the per
        // element weight w = stress weight |X|^2 every component is a line
        // constant times W = sum w, Wz = sum w kz or Wzz = sum w kz^2.  The
        // Hermitian partner (-j, -ky, -kz) of an interior column (j not 0 or
        // K1/2) counts too, with a Nyquist index (K/2) kept at +K/2 (the
        // reference folds it so), which only the off-diagonal terms see: its
        // kz sum is Wzp = -Wz + 2 Wn, Wn the Nyquist element's w kz.
        double W = 0.0, Wz = 0.0, Wzz = 0.0, Wn = 0.0;
#pragma unroll
        for (int r = 0; r < RB; r++) {
            const int m1 = t + N2 * r;
            if (m1 < N1) {
                T2 y[N2];
                stage_b<T, N1, N2, 1>(m1, ex0 + ll * EXS, y);
                T2 u[N2];
#pragma unroll
                for (int p = 0; p < N2; p++) {
                    const int l = m1 + N1 * D2::perm(p);
                    const T2 h = y[p], f = thv[r][p];
                    const double w = (double)(f.y * (h.x * h.x + h.y * h.y));
                    const double wl = w * (double)((2 * l > ip.K[2]) ? l - ip.K[2] : l);
                    W += w;
                    Wz += wl;
                    Wzz += wl * (double)((2 * l > ip.K[2]) ? l - ip.K[2] : l);
                    if (2 * l == ip.K[2]) Wn += wl;
                    T2 g;
                    g.x = f.x * h.x;
                    g.y = f.x * h.y;
                    u[D2::perm(p)] = g;  // natural order m2 for the inverse
                }
                // inverse over m2: u[q] = V[m1][n2 = perm(q)], then the
                // twiddle exp(-2 pi i n2 m1 / N) and the exchange
                RegDft<T, T2, N2, -1>::template run<0, 1>(u);
#pragma unroll
                for (int q = 0; q < N2; q++) {
                    const int n2 = RegDft<T, T2, N2, -1>::perm(q);
                    T2 v = u[q];
                    if (n2 != 0) {
                        T2 w = stw[n2 * m1];  // n2 m1 < N
                        w.y = -w.y;
                        v = cmul(v, w);
                    }
                    ex1[ll * EXS + m1 * N2 + n2] = v;
                }
            }
        }
        if (on) {
            const int j = L % lm.na_pad, k = L / lm.na_pad;
            const int kk = (2 * k > ip.K[1]) ? k - ip.K[1] : k;
            const int kkp = (2 * k == ip.K[1]) ? kk : -kk;
            const bool partner = !(j == 0 || 2 * j == ip.K[0]);
            const double mx = j * ip.rcell[0], my = kk * ip.rcell[4], myp = kkp * ip.rcell[4], r8 = ip.rcell[8];
            const double cp = partner ? 2.0 : 1.0, Wzp = -Wz + 2.0 * Wn;
            acc[0] = cp * mx * mx * W;
            acc[1] = mx * my * W + (partner ? -mx * myp * W : 0.0);
            acc[2] = mx * r8 * Wz + (partner ? -mx * r8 * Wzp : 0.0);
            acc[3] = cp * my * my * W;
            acc[4] = my * r8 * Wz + (partner ? myp * r8 * Wzp : 0.0);
            acc[5] = cp * r8 * r8 * Wzz;
        }
    }
    __syncthreads();
    if (t < N2) {
        T2 z[N1];
#pragma unroll
        for (int m1 = 0; m1 < N1; m1++) z[m1] = ex1[ll * EXS + m1 * N2 + t];
        RegDft<T, T2, N1, -1>::template run<0, 1>(z);
        if (on)
#pragma unroll
            for (int p = 0; p < N1; p++)
                a.data[base + (size_t)(t + N2 * RegDft<T, T2, N1, -1>::perm(p)) * lm.se] = z[p];
    }
    block_sum<6>(acc, lred, a.red + (size_t)bid * kReduceSlots + R_RS11);
}

// Lines of the x passes: LPB pairs of real lines (2P, 2P+1) per workgroup,
// moved between the grid and LDS in whole rows (consecutive threads on
// consecutive elements), UN values per thread all in flight together.
template <int LPB, int N, int NT>
constexpr int x_per_thread()
{
    return (LPB * N + NT - 1) / NT;
}

// ---- x axis, forward: two real lines per complex line, split into half
// spectra --------------------------------------------------------------------
template <typename T, int N1, int N2, int LPB>
__device__ __forceinline__ void lines_r2c_body(const R2CArgs<T> &a, int bid)
{
    using T2 = typename C2T<T>::type;
    constexpr int N = N1 * N2, EXS = N + 1, RB = (N1 + N2 - 1) / N2, K1h = N / 2 + 1;
    constexpr int NT = (LPB * N2 + 63) / 64 * 64, UN = x_per_thread<LPB, N, NT>();
    extern __shared__ __align__(16) unsigned char smem[];
    T2 *ex0 = (T2 *)smem, *ex1 = ex0 + LPB * EXS, *stw = ex1 + LPB * EXS;
    T *__restrict__ Q = a.Q;
    const int nreal = a.nreal_lines, npairs = (nreal + 1) / 2;
    const int P0 = bid * LPB, np = min(LPB, npairs - P0);
    {
        // with clear, each value read is zeroed for the next spreading
        T2 z[UN];
#pragma unroll
        for (int u = 0; u < UN; u++) {
            const int idx = threadIdx.x + u * NT;
            const int l2 = idx / N, e = idx - l2 * N;
            const size_t ra = 2 * (size_t)(P0 + l2), rb = ra + 1;
            z[u] = zero2<T2>();
            if (l2 < np) {
                z[u].x = Q[ra * N + e];
                if (rb < (size_t)nreal) z[u].y = Q[rb * N + e];
            }
        }
#pragma unroll
        for (int u = 0; u < UN; u++) {
            const int idx = threadIdx.x + u * NT;
            const int l2 = idx / N, e = idx - l2 * N;
            const size_t ra = 2 * (size_t)(P0 + l2), rb = ra + 1;
            if (l2 < np) {
                ex0[l2 * EXS + e] = z[u];
                if (a.clear) {
                    Q[ra * N + e] = (T)0;
                    if (rb < (size_t)nreal) Q[rb * N + e] = (T)0;
                }
            }
        }
    }
    for (int k = threadIdx.x; k < N; k += NT) stw[k] = a.tw[k];
    __syncthreads();
    const int ll = threadIdx.x % LPB, t = threadIdx.x / LPB;
    const bool on = t < N2 && ll < np;
    if (on) {
        T2 x[N1];
#pragma unroll
        for (int n1 = 0; n1 < N1; n1++) x[n1] = ex0[ll * EXS + t + N2 * n1];
        stage_a<T, N1, N2, 1>(x, t, ex1 + ll * EXS, stw);
    }
    __syncthreads();
    if (on) {
#pragma unroll
        for (int r = 0; r < RB; r++) {
            const int m1 = t + N2 * r;
            if (m1 < N1) {
                T2 y[N2];
                stage_b<T, N1, N2, 1>(m1, ex1 + ll * EXS, y);
#pragma unroll
                for (int p = 0; p < N2; p++) ex0[ll * EXS + m1 + N1 * RegDft<T, T2, N2, 1>::perm(p)] = y[p];
            }
        }
    }
    __syncthreads();
    T2 *__restrict__ H = a.H;
    const int hs = a.hs;
    constexpr int UH = (LPB * K1h + NT - 1) / NT;
#pragma unroll
    for (int u = 0; u < UH; u++) {
        const int idx = threadIdx.x + u * NT;
        const int l2 = idx / K1h, k = idx - l2 * K1h;
        if (l2 < np) {
            const size_t ra = 2 * (size_t)(P0 + l2), rb = ra + 1;
            const T2 z = ex0[l2 * EXS + k];
            const T2 zc = ex0[l2 * EXS + (k == 0 ? 0 : N - k)];
            T2 A, B;
            A.x = (T)0.5 * (z.x + zc.x);
            A.y = (T)0.5 * (z.y - zc.y);
            B.x = (T)0.5 * (z.y + zc.y);
            B.y = (T)-0.5 * (z.x - zc.x);
            H[ra * hs + k] = A;
            if (rb < (size_t)nreal) H[rb * hs + k] = B;
        }
    }
}

// ---- x axis, inverse: the full Hermitian line of two half spectra, sign -1,
// real parts to line 2P and imaginary parts to line 2P+1 ----------------------
template <typename T, int N1, int N2, int LPB>
__device__ __forceinline__ void lines_c2r_body(const C2RArgs<T> &a, int bid)
{
    using T2 = typename C2T<T>::type;
    constexpr int N = N1 * N2, EXS = N + 1, RB = (N1 + N2 - 1) / N2, K1h = N / 2 + 1;
    constexpr int NT = (LPB * N2 + 63) / 64 * 64, UH = (LPB * 2 * K1h + NT - 1) / NT,
                  UN = x_per_thread<LPB, N, NT>();
    extern __shared__ __align__(16) unsigned char smem[];
    // hb: the two half spectra of each pair (2 K1h), later the output lines
    T2 *ex = (T2 *)smem, *stw = ex + LPB * EXS, *hb = stw + N;
    constexpr int HBS = 2 * K1h > N ? 2 * K1h : N;
    const T2 *__restrict__ H = a.H;
    T *__restrict__ Q = a.Q;
    const int nreal = a.nreal_lines, npairs = (nreal + 1) / 2, hs = a.hs;
    const int P0 = bid * LPB, np = min(LPB, npairs - P0);
    {
        T2 v[UH];
#pragma unroll
        for (int u = 0; u < UH; u++) {
            const int idx = threadIdx.x + u * NT;
            const int l2 = idx / (2 * K1h), w = idx - l2 * (2 * K1h);
            const size_t row = 2 * (size_t)(P0 + l2) + (w >= K1h ? 1 : 0);
            v[u] = zero2<T2>();
            if (l2 < np && row < (size_t)nreal) v[u] = H[row * hs + (w >= K1h ? w - K1h : w)];
        }
#pragma unroll
        for (int u = 0; u < UH; u++) {
            const int idx = threadIdx.x + u * NT;
            const int l2 = idx / (2 * K1h), w = idx - l2 * (2 * K1h);
            if (l2 < LPB) hb[l2 * HBS + w] = v[u];
        }
    }
    for (int k = threadIdx.x; k < N; k += NT) stw[k] = a.tw[k];
    __syncthreads();
    const int ll = threadIdx.x % LPB, t = threadIdx.x / LPB;
    const bool on = t < N2 && ll < np;
    if (on) {
        T2 x[N1];
        const T2 *hl = hb + ll * HBS;
#pragma unroll
        for (int n1 = 0; n1 < N1; n1++) {
            const int k = t + N2 * n1;
            const int kh = k < K1h ? k : N - k;
            const T2 A = hl[kh], B = hl[K1h + kh];
            T2 z;
            if (k < K1h) {
                z.x = A.x - B.y;
                z.y = A.y + B.x;
            } else {
                z.x = A.x + B.y;
                z.y = -A.y + B.x;
            }
            x[n1] = z;
        }
        stage_a<T, N1, N2, -1>(x, t, ex + ll * EXS, stw);
    }
    __syncthreads();
    if (on) {
#pragma unroll
        for (int r = 0; r < RB; r++) {
            const int m1 = t + N2 * r;
            if (m1 < N1) {
                T2 y[N2];
                stage_b<T, N1, N2, -1>(m1, ex + ll * EXS, y);
#pragma unroll
                for (int p = 0; p < N2; p++) hb[ll * HBS + m1 + N1 * RegDft<T, T2, N2, -1>::perm(p)] = y[p];
            }
        }
    }
    __syncthreads();
#pragma unroll
    for (int u = 0; u < UN; u++) {
        const int idx = threadIdx.x + u * NT;
        const int l2 = idx / N, e = idx - l2 * N;
        if (l2 < np) {
            const size_t ra = 2 * (size_t)(P0 + l2), rb = ra + 1;
            const T2 z = hb[l2 * HBS + e];
            Q[ra * N + e] = z.x;
            if (rb < (size_t)nreal) Q[rb * N + e] = z.y;
        }
    }
}

// Single-system kernels (arguments by value) and batched ones (one argument
// struct per system, see batch_find).
#define DLPX_LINES_KERNELS(NAME, ARGS, BODY, TPARAMS, TARGS)                                                  \
    template <TPARAMS>                                                                                          \
    __global__ void __launch_bounds__((LPB * N2 + 63) / 64 * 64) NAME(ARGS<T> a)                                \
    {                                                                                                           \
        BODY<TARGS>(a, blockIdx.x);                                                                             \
    }                                                                                                           \
    template <TPARAMS>                                                                                          \
    __global__ void __launch_bounds__((LPB * N2 + 63) / 64 * 64)                                                \
        NAME##_batch(const ARGS<T> *__restrict__ av, const int32_t *__restrict__ boff, int nsys)               \
    {                                                                                                           \
        const int lb = blockIdx.x;                                                                              \
        if (lb >= boff[nsys]) return;                                                                           \
        const int s = batch_find(boff, nsys, lb);                                                               \
        BODY<TARGS>(av[s], lb - boff[s]);                                                                       \
    }
#define DLPX_COMMA ,
DLPX_LINES_KERNELS(k_lines_c2c, C2CArgs, lines_c2c_body,
                   typename T DLPX_COMMA int N1 DLPX_COMMA int N2 DLPX_COMMA int LPB DLPX_COMMA int SGN,
                   T DLPX_COMMA N1 DLPX_COMMA N2 DLPX_COMMA LPB DLPX_COMMA SGN)
DLPX_LINES_KERNELS(k_lines_zconv, ZConvArgs, lines_zconv_body,
                   typename T DLPX_COMMA int N1 DLPX_COMMA int N2 DLPX_COMMA int LPB,
                   T DLPX_COMMA N1 DLPX_COMMA N2 DLPX_COMMA LPB)
DLPX_LINES_KERNELS(k_lines_r2c, R2CArgs, lines_r2c_body,
                   typename T DLPX_COMMA int N1 DLPX_COMMA int N2 DLPX_COMMA int LPB,
                   T DLPX_COMMA N1 DLPX_COMMA N2 DLPX_COMMA LPB)
DLPX_LINES_KERNELS(k_lines_c2r, C2RArgs, lines_c2r_body,
                   typename T DLPX_COMMA int N1 DLPX_COMMA int N2 DLPX_COMMA int LPB,
                   T DLPX_COMMA N1 DLPX_COMMA N2 DLPX_COMMA LPB)
#undef DLPX_COMMA
#undef DLPX_LINES_KERNELS

// ---- host side -------------------------------------------------------------
bool lines_supported(int n)
{
    switch (n) {
#define DLPX_CASE(N, A, B) case N:
        DLPX_LINE_LENGTHS(DLPX_CASE)
#undef DLPX_CASE
        return true;
        default: return false;
    }
}

static int line_split2(int n)
{
    switch (n) {
#define DLPX_CASE(N, A, B) case N: return B;
        DLPX_LINE_LENGTHS(DLPX_CASE)
#undef DLPX_CASE
        default: return 0;
    }
}

template <typename T>
int lines_lpb()
{
    // 128-byte runs of the strided y and z lines: 16 columns of FP32
    // complex, 8 of FP64
    return sizeof(T) == 4 ? 16 : 8;
}
template int lines_lpb<float>();
template int lines_lpb<double>();

template <typename T>
static int threads_for(int n)
{
    return (lines_lpb<T>() * line_split2(n) + 63) / 64 * 64;
}

template <typename T>
size_t lines_lds(int kind, int n)
{
    const size_t e = sizeof(typename C2T<T>::type), lpb = lines_lpb<T>(), rows = lpb * (n + 1);
    switch (kind) {
        case 0: return e * (rows + n);                      // c2c
        case 1: return kConvRedBytes + e * (2 * rows + n);  // z pass
        case 2: return e * (2 * rows + n);                  // r2c
        default: return e * (rows + n + lpb * std::max(2 * (n / 2 + 1), n));  // c2r
    }
}
template size_t lines_lds<float>(int, int);
template size_t lines_lds<double>(int, int);

template <typename T>
int lines_c2c(Ctx *c, hipStream_t s, const C2CArgs<T> &a, int nb)
{
    constexpr int LPB = sizeof(T) == 4 ? 16 : 8;
    const int n = a.pl.n, nt = threads_for<T>(n);
    const size_t lds = lines_lds<T>(0, n);
    switch (n) {
#define DLPX_CASE(N, A, B)                                                                                        \
    case N:                                                                                                       \
        if (a.sign > 0) {                                                                                         \
            DLPX_LDS_MAX(c, (k_lines_c2c<T, A, B, LPB, 1>));                                                      \
            hipLaunchKernelGGL((k_lines_c2c<T, A, B, LPB, 1>), dim3(nb), dim3(nt), lds, s, a);                    \
        } else {                                                                                                  \
            DLPX_LDS_MAX(c, (k_lines_c2c<T, A, B, LPB, -1>));                                                     \
            hipLaunchKernelGGL((k_lines_c2c<T, A, B, LPB, -1>), dim3(nb), dim3(nt), lds, s, a);                   \
        }                                                                                                         \
        break;
        DLPX_LINE_LENGTHS(DLPX_CASE)
#undef DLPX_CASE
        default: return fail(c, DLPX_ERR_STATE, "no line kernel for this FFT length");
    }
    DLPX_HIP(c, hipGetLastError());
    return 0;
}

template <typename T>
int lines_zconv(Ctx *c, hipStream_t s, const ZConvArgs<T> &a, int nb)
{
    constexpr int LPB = sizeof(T) == 4 ? 16 : 8;
    const int n = a.pl.n, nt = threads_for<T>(n);
    const size_t lds = lines_lds<T>(1, n);
    switch (n) {
#define DLPX_CASE(N, A, B)                                                                                        \
    case N:                                                                                                       \
        DLPX_LDS_MAX(c, (k_lines_zconv<T, A, B, LPB>));                                                           \
        hipLaunchKernelGGL((k_lines_zconv<T, A, B, LPB>), dim3(nb), dim3(nt), lds, s, a);                         \
        break;
        DLPX_LINE_LENGTHS(DLPX_CASE)
#undef DLPX_CASE
        default: return fail(c, DLPX_ERR_STATE, "no line kernel for this FFT length");
    }
    DLPX_HIP(c, hipGetLastError());
    return 0;
}

template <typename T>
int lines_x(Ctx *c, hipStream_t s, const R2CArgs<T> *fwd, const C2RArgs<T> *inv, int nb)
{
    constexpr int LPB = sizeof(T) == 4 ? 16 : 8;
    const int n = fwd ? fwd->pl.n : inv->pl.n, nt = threads_for<T>(n);
    const size_t lds = lines_lds<T>(fwd ? 2 : 3, n);
    switch (n) {
#define DLPX_CASE(N, A, B)                                                                                        \
    case N:                                                                                                       \
        if (fwd) {                                                                                                \
            DLPX_LDS_MAX(c, (k_lines_r2c<T, A, B, LPB>));                                                         \
            hipLaunchKernelGGL((k_lines_r2c<T, A, B, LPB>), dim3(nb), dim3(nt), lds, s, *fwd);                    \
        } else {                                                                                                  \
            DLPX_LDS_MAX(c, (k_lines_c2r<T, A, B, LPB>));                                                         \
            hipLaunchKernelGGL((k_lines_c2r<T, A, B, LPB>), dim3(nb), dim3(nt), lds, s, *inv);                    \
        }                                                                                                         \
        break;
        DLPX_LINE_LENGTHS(DLPX_CASE)
#undef DLPX_CASE
        default: return fail(c, DLPX_ERR_STATE, "no line kernel for this FFT length");
    }
    DLPX_HIP(c, hipGetLastError());
    return 0;
}

template <typename T>
int lines_batch(Ctx *c, hipStream_t s, int kind, int n, const BatchLaunch &bl)
{
    constexpr int LPB = sizeof(T) == 4 ? 16 : 8;
    const int nt = threads_for<T>(n);
    const dim3 g(bl.total), b(nt);
#define DLPX_BL(KERNEL, ARGS, ...)                                                                                \
    do {                                                                                                          \
        DLPX_LDS_MAX(c, (KERNEL<__VA_ARGS__>));                                                                   \
        hipLaunchKernelGGL((KERNEL<__VA_ARGS__>), g, b, bl.lds, s, (const ARGS<T> *)bl.args, bl.boff, bl.nsys); \
    } while (0)
    switch (n) {
#define DLPX_CASE(N, A, B)                                                                                        \
    case N:                                                                                                       \
        switch (kind) {                                                                                           \
            case 0: DLPX_BL(k_lines_r2c_batch, R2CArgs, T, A, B, LPB); break;                                     \
            case 1: DLPX_BL(k_lines_c2c_batch, C2CArgs, T, A, B, LPB, 1); break;                                  \
            case 2: DLPX_BL(k_lines_zconv_batch, ZConvArgs, T, A, B, LPB); break;                                 \
            case 3: DLPX_BL(k_lines_c2c_batch, C2CArgs, T, A, B, LPB, -1); break;                                 \
            default: DLPX_BL(k_lines_c2r_batch, C2RArgs, T, A, B, LPB); break;                                    \
        }                                                                                                         \
        break;
        DLPX_LINE_LENGTHS(DLPX_CASE)
#undef DLPX_CASE
        default: return fail(c, DLPX_ERR_STATE, "no line kernel for this FFT length");
    }
#undef DLPX_BL
    DLPX_HIP(c, hipGetLastError());
    return 0;
}

template int lines_batch<float>(Ctx *, hipStream_t, int, int, const BatchLaunch &);
template int lines_batch<double>(Ctx *, hipStream_t, int, int, const BatchLaunch &);
template int lines_c2c<float>(Ctx *, hipStream_t, const C2CArgs<float> &, int);
template int lines_c2c<double>(Ctx *, hipStream_t, const C2CArgs<double> &, int);
template int lines_zconv<float>(Ctx *, hipStream_t, const ZConvArgs<float> &, int);
template int lines_zconv<double>(Ctx *, hipStream_t, const ZConvArgs<double> &, int);
template int lines_x<float>(Ctx *, hipStream_t, const R2CArgs<float> *, const C2RArgs<float> *, int);
template int lines_x<double>(Ctx *, hipStream_t, const R2CArgs<double> *, const C2RArgs<double> *, int);

}  // namespace dlpx
