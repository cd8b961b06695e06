// SPME reciprocal space (ewald_spole.F90:244-477, ewald_general.F90:517-869,
// bspline.F90:192-306, ewald_spole.F90:1257-1386) on the GPU.
//
//  k_spread     B-splines of 64 atoms per block go to LDS; each wavefront then
//               scatters one atom's S^3 charge weights, lanes over the (x, y)
//               plane so one atomic instruction touches S rows of S contiguous
//               grid points.
//  k_influence  Q^ <- theta(m) Q^ on the half spectrum with the spherical
//               cutoff 0.525 min_d K_d w*_d, and the reciprocal stress.
//  k_gather     same LDS staging; lanes over (x, y) read the potential grid,
//               wave reduction gives the atom's energy and force.
//  k_finish     subtracts the mean reciprocal force (ewald_general.F90:857-864),
//               scales by 2*scale and adds into the total force.
// FP64 or FP32 (mixed) arithmetic; sums across lanes and atoms in FP64.
#include <cmath>
#include <cstring>
#include <vector>

#include "common.h"
#include "device_util.h"

namespace dlpx {

struct SpmeParams {
    int K[3];
    int K1h;
    int S;
    double kr[3];           // K as real
    double rcell[9];        // inverse cell (numerics.F90 invert)
    double rk[3];           // Matmul(reshape(rcell,[3,3]), K) (ewald_general.F90:770-771)
};

// bspline_splines_gen (bspline.F90:192-306) for one coordinate: order-S
// values M[t] and first derivatives D[t] at the S grid points
// floor(u) - S + 1 + t, t = 0..S-1, evaluated with the reference's recursion
// (S even in 4..16; the reference forces an even order, control.F90:1651-1675).
template <typename R, bool DERIV, int S>
__device__ __forceinline__ void bspline(R w, R *M, R *D)
{
    R zd[S + 2];
#pragma unroll
    for (int j = 0; j < S + 2; j++) zd[j] = (R)0;
    // 1-based indices as in the reference: zd[1..S], zd[S+1] == 0
    zd[S] = w;
    zd[S - 1] = (R)1 - w;
#pragma unroll
    for (int k = S - 2; k >= 3; k--) {
        const R k_r = (R)(S - k + 1);
        const R km1_rr = (R)1 / (R)(S - k);
#pragma unroll
        for (int j = k; j <= S - 1; j++) {
            const R pt = w + (R)(S - j);
            zd[j] = (pt * zd[j] + (k_r - pt) * zd[j + 1]) * km1_rr;
        }
        zd[S] = zd[S] * w * km1_rr;
    }
    // l = 2 stage: raise to order S-1 (second derivatives are not needed)
    {
        const R k_r = (R)(S - 1);
        const R km1_rr = (R)1 / (R)(S - 2);
#pragma unroll
        for (int j = 1; j <= S - 1; j++) {
            const R pt = w + (R)(S - j);
            zd[j] = (pt * zd[j] + (k_r - pt) * zd[j + 1]) * km1_rr;
        }
        zd[S] = zd[S] * w * km1_rr;
    }
    // l = 1 stage: first derivatives from order S-1, then raise to order S
    {
        const R k_r = (R)S;
        const R km1_rr = (R)1 / (R)(S - 1);
#pragma unroll
        for (int j = 1; j <= S - 1; j++) {
            if (DERIV) D[j - 1] = zd[j] - zd[j + 1];
            const R pt = w + (R)(S - j);
            zd[j] = (pt * zd[j] + (k_r - pt) * zd[j + 1]) * km1_rr;
        }
        if (DERIV) D[S - 1] = zd[S];
        zd[S] = zd[S] * w * km1_rr;
    }
#pragma unroll
    for (int j = 1; j <= S; j++) M[j - 1] = zd[j];
}

// Fractional grid coordinate of one axis, folded into [0, K) so an atom that
// drifted out of the cell is represented by its periodic image
// (ewald_spole.F90:364-387).
__device__ __forceinline__ void grid_coord(const SpmeParams &sp, int d, double x, double y, double z, int &idx,
                                           double &w)
{
    double u = sp.kr[d] * (sp.rcell[d] * x + sp.rcell[d + 3] * y + sp.rcell[d + 6] * z + 0.5);
    if (u < 0.0 || u >= sp.kr[d]) u -= sp.kr[d] * floor(u / sp.kr[d]);
    int i = (int)u;
    if (i >= sp.K[d]) i = sp.K[d] - 1;  // u rounded up to K after folding
    if (!(i >= 0)) i = 0;               // non-finite coordinates stay inside the grid
    idx = i;
    w = u - (double)i;
}

constexpr int kAtomsPerBlock = 64;

template <typename R, int S>
__global__ void __launch_bounds__(256) k_spread(const double4 *__restrict__ pos, int n, SpmeParams sp,
                                                R *__restrict__ Q)
{
    __shared__ R sM[3][kAtomsPerBlock][S];
    __shared__ int sI[3][kAtomsPerBlock];
    __shared__ R sQ[kAtomsPerBlock];
    const int a0 = blockIdx.x * kAtomsPerBlock;
    const int t = threadIdx.x;
    if (t < 3 * kAtomsPerBlock) {
        const int d = t / kAtomsPerBlock, a = t % kAtomsPerBlock;
        const int i = a0 + a;
        if (i < n) {
            const double4 p = pos[i];
            int idx;
            double w;
            grid_coord(sp, d, p.x, p.y, p.z, idx, w);
            R M[S];
            bspline<R, false, S>((R)w, M, nullptr);
#pragma unroll
            for (int k = 0; k < S; k++) sM[d][a][k] = M[k];
            sI[d][a] = idx;
            if (d == 0) sQ[a] = (R)p.w;
        }
    }
    __syncthreads();
    const int wid = t >> 6, lane = t & 63;
    const int na = min(kAtomsPerBlock, n - a0);
    for (int a = wid; a < na; a += 4) {
        const R q = sQ[a];
        if (fabs((double)q) <= kZeroPlus) continue;
        const int bx = sI[0][a] - S + 1 + sp.K[0], by = sI[1][a] - S + 1 + sp.K[1], bz = sI[2][a] - S + 1 + sp.K[2];
        for (int p = lane; p < S * S; p += 64) {
            const int tx = p % S, ty = p / S;
            const int gx = (bx + tx) % sp.K[0], gy = (by + ty) % sp.K[1];
            const R mxy = sM[0][a][tx] * sM[1][a][ty];
#pragma unroll
            for (int tz = 0; tz < S; tz++) {
                const int gz = (bz + tz) % sp.K[2];
                unsafeAtomicAdd(&Q[((size_t)gz * sp.K[1] + gy) * sp.K[0] + gx], q * sM[2][a][tz] * mxy);
            }
        }
    }
}

// Unwrapped fractional grid coordinate of one axis (no folding): the base
// grid point floor(u) - S + 1 may be negative or >= K; callers fold indices.
__device__ __forceinline__ double grid_u(const SpmeParams &sp, int d, double x, double y, double z)
{
    return sp.kr[d] * (sp.rcell[d] * x + sp.rcell[d + 3] * y + sp.rcell[d + 6] * z + 0.5);
}

__device__ __forceinline__ int wrapk(int v, int K)
{
    v %= K;
    return v < 0 ? v + K : v;
}

constexpr int kSlabBytes = 32768;

// Charge spreading by link cell (one workgroup per cell of the sorted atom
// order): the cell's atoms touch a compact window of grid points, which is
// accumulated in LDS in z-slabs with LDS atomics and then added to the global
// grid with one atomic per window point, instead of S^3 global atomics per
// atom.  A window too large for LDS (very large cells, or atoms far outside
// their cell) falls back to per-atom global atomics for that cell.
template <typename R, int S>
__global__ void __launch_bounds__(256) k_spread_cells(const double4 *__restrict__ pos,
                                                      const int32_t *__restrict__ cell_start, SpmeParams sp,
                                                      R *__restrict__ Q)
{
    constexpr int kCap = kSlabBytes / (int)sizeof(R);
    __shared__ R slab[kCap];
    __shared__ R sM[3][kAtomsPerBlock][S];
    __shared__ int sB[3][kAtomsPerBlock];
    __shared__ R sQ[kAtomsPerBlock];
    __shared__ int wlo[3], whi[3];
    const int c = blockIdx.x, t = threadIdx.x, wid = t >> 6, lane = t & 63;
    const int s0 = cell_start[c], n = cell_start[c + 1] - s0;
    if (n <= 0) return;
    if (t < 3) {
        wlo[t] = 0x7fffffff;
        whi[t] = -0x7fffffff;
    }
    __syncthreads();
    for (int k = t; k < n; k += blockDim.x) {
        const double4 p = pos[s0 + k];
        for (int d = 0; d < 3; d++) {
            const int b = (int)floor(grid_u(sp, d, p.x, p.y, p.z)) - S + 1;
            atomicMin(&wlo[d], b);
            atomicMax(&whi[d], b);
        }
    }
    __syncthreads();
    const int lo0 = wlo[0], lo1 = wlo[1], lo2 = wlo[2];
    // window extents in 64-bit so that non-finite coordinates cannot overflow
    const long ex = (long)whi[0] - lo0 + S, ey = (long)whi[1] - lo1 + S, ez = (long)whi[2] - lo2 + S;
    const bool tile = ex <= sp.K[0] && ey <= sp.K[1] && ez <= sp.K[2] && ex * ey <= kCap;
    const int Wx = tile ? (int)ex : 0, Wy = tile ? (int)ey : 0, Wz = tile ? (int)ez : 1;
    const int Dz = tile ? min(Wz, kCap / (Wx * Wy)) : Wz;
    for (int z0 = 0; z0 < Wz; z0 += Dz) {
        if (tile) {
            for (int k = t; k < Wx * Wy * Dz; k += blockDim.x) slab[k] = (R)0;
        }
        for (int a0 = 0; a0 < n; a0 += kAtomsPerBlock) {
            __syncthreads();
            if (t < 3 * kAtomsPerBlock) {
                const int d = t / kAtomsPerBlock, a = t % kAtomsPerBlock;
                if (a0 + a < n) {
                    const double4 p = pos[s0 + a0 + a];
                    const double u = grid_u(sp, d, p.x, p.y, p.z);
                    const double fl = floor(u);
                    R M[S];
                    bspline<R, false, S>((R)(u - fl), M, nullptr);
#pragma unroll
                    for (int k = 0; k < S; k++) sM[d][a][k] = M[k];
                    sB[d][a] = (int)fl - S + 1;
                    if (d == 0) sQ[a] = (R)p.w;
                }
            }
            __syncthreads();
            const int na = min(kAtomsPerBlock, n - a0);
            for (int a = wid; a < na; a += 4) {
                const R q = sQ[a];
                if (fabs((double)q) <= kZeroPlus) continue;
                const int bx = sB[0][a], by = sB[1][a], bz = sB[2][a];
                for (int pp = lane; pp < S * S; pp += 64) {
                    const int tx = pp % S, ty = pp / S;
                    const R mxy = sM[0][a][tx] * sM[1][a][ty];
                    if (tile) {
                        const int lxy = (by + ty - lo1) * Wx + (bx + tx - lo0);
#pragma unroll
                        for (int tz = 0; tz < S; tz++) {
                            const int lz = bz + tz - lo2 - z0;
                            const R v = q * sM[2][a][tz] * mxy;
                            // single precision: contributions below 1e-30 (spline tails)
                            // are dropped rather than added as denormals
                            if (lz >= 0 && lz < Dz && (sizeof(R) == 8 || fabs(v) >= (R)1e-30))
                                atomicAdd(&slab[lz * Wx * Wy + lxy], v);
                        }
                    } else {
                        const int gx = wrapk(bx + tx, sp.K[0]), gy = wrapk(by + ty, sp.K[1]);
#pragma unroll
                        for (int tz = 0; tz < S; tz++) {
                            const int gz = wrapk(bz + tz, sp.K[2]);
                            unsafeAtomicAdd(&Q[((size_t)gz * sp.K[1] + gy) * sp.K[0] + gx], q * sM[2][a][tz] * mxy);
                        }
                    }
                }
            }
        }
        __syncthreads();
        if (!tile) break;  // the direct path did every z at once
        const int dz = min(Dz, Wz - z0);
        for (int k = t; k < Wx * Wy * dz; k += blockDim.x) {
            const R v = slab[k];
            if (v == (R)0) continue;
            const int lx = k % Wx, ly = (k / Wx) % Wy, lz = k / (Wx * Wy);
            const int gx = wrapk(lo0 + lx, sp.K[0]), gy = wrapk(lo1 + ly, sp.K[1]), gz = wrapk(lo2 + z0 + lz, sp.K[2]);
            unsafeAtomicAdd(&Q[((size_t)gz * sp.K[1] + gy) * sp.K[0] + gx], v);
        }
    }
}

// Influence function and reciprocal stress over the half spectrum
// (x index 0..K1/2), ewald_spole.F90:1296-1380.
struct InflParams {
    int K[3], K1h;
    double rcell[9];
    double cut2, test_fac, rcf;  // cut_off^2, (1e-6/rcf)^2, pi/alpha
    int nblocks;
};

template <typename R>
__global__ void __launch_bounds__(256) k_influence(typename std::conditional<sizeof(R) == 8, double2, float2>::type *H,
                                                   const double *__restrict__ n2, InflParams ip, double *red)
{
    __shared__ double lred[6 * 4];
    const double sqrpi = 1.7724538509055159;
    double acc[6] = {0, 0, 0, 0, 0, 0};
    const size_t total = (size_t)ip.K1h * ip.K[1] * ip.K[2];
    const double *n2x = n2, *n2y = n2 + ip.K[0], *n2z = n2 + ip.K[0] + ip.K[1];
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (size_t)gridDim.x * blockDim.x) {
        const int j = (int)(e % ip.K1h);
        const int k = (int)((e / ip.K1h) % ip.K[1]);
        const int l = (int)(e / ((size_t)ip.K1h * ip.K[1]));
        const int jj = j;
        const int kk = (2 * k > ip.K[1]) ? k - ip.K[1] : k;
        const int ll = (2 * l > ip.K[2]) ? l - ip.K[2] : l;
        double r3[3] = {ll * ip.rcell[2], ll * ip.rcell[5], ll * ip.rcell[8]};
        double r2[3] = {r3[0] + kk * ip.rcell[1], r3[1] + kk * ip.rcell[4], r3[2] + kk * ip.rcell[7]};
        double rp[3] = {r2[0] + jj * ip.rcell[0], r2[1] + jj * ip.rcell[3], r2[2] + jj * ip.rcell[6]};
        const double bb1 = (n2z[l] * n2y[k]) * n2x[j];
        const double k2 = rp[0] * rp[0] + rp[1] * rp[1] + rp[2] * rp[2];
        auto h = H[e];
        if (k2 <= ip.cut2 && k2 > ip.test_fac) {
            const double fp = sqrt(k2) * ip.rcf;
            const double fac = bb1 * exp(-(fp * fp)) / (sqrpi * (fp * fp));
            const double hr = (double)h.x, hi = (double)h.y;
            const double pv = fac * (hr * hr + hi * hi) * (-2.0 * ((1.0 + fp * fp) / k2));
            const double wgt = (j == 0 || 2 * j == ip.K[0]) ? 1.0 : 2.0;  // Hermitian partner
            acc[0] += wgt * rp[0] * rp[0] * pv;
            acc[1] += wgt * rp[0] * rp[1] * pv;
            acc[2] += wgt * rp[0] * rp[2] * pv;
            acc[3] += wgt * rp[1] * rp[1] * pv;
            acc[4] += wgt * rp[1] * rp[2] * pv;
            acc[5] += wgt * rp[2] * rp[2] * pv;
            h.x = (R)(fac * hr);
            h.y = (R)(fac * hi);
        } else {
            h.x = (R)0;
            h.y = (R)0;
        }
        H[e] = h;
    }
    block_sum<6>(acc, lred, red + (size_t)blockIdx.x * kReduceSlots + R_RS11);
}

template <typename R, int S>
__global__ void __launch_bounds__(256) k_gather(const double4 *__restrict__ pos, int n, SpmeParams sp,
                                                const R *__restrict__ Phi, double4 *__restrict__ frc_recip,
                                                double *red)
{
    __shared__ R sM[3][kAtomsPerBlock][S];
    __shared__ R sD[3][kAtomsPerBlock][S];
    __shared__ int sI[3][kAtomsPerBlock];
    __shared__ double sQ[kAtomsPerBlock];
    __shared__ double lred[4 * 4];
    const int a0 = blockIdx.x * kAtomsPerBlock;
    const int t = threadIdx.x;
    if (t < 3 * kAtomsPerBlock) {
        const int d = t / kAtomsPerBlock, a = t % kAtomsPerBlock;
        const int i = a0 + a;
        if (i < n) {
            const double4 p = pos[i];
            int idx;
            double w;
            grid_coord(sp, d, p.x, p.y, p.z, idx, w);
            R M[S], D[S];
            bspline<R, true, S>((R)w, M, D);
#pragma unroll
            for (int k = 0; k < S; k++) {
                sM[d][a][k] = M[k];
                sD[d][a][k] = D[k];
            }
            sI[d][a] = idx;
            if (d == 0) sQ[a] = p.w;
        }
    }
    __syncthreads();
    const int wid = t >> 6, lane = t & 63;
    const int na = min(kAtomsPerBlock, n - a0);
    double acc[4] = {0, 0, 0, 0};  // energy, fx, fy, fz (unscaled, summed over atoms)
    for (int a = wid; a < na; a += 4) {
        const double q = sQ[a];
        double e = 0.0, fx = 0.0, fy = 0.0, fz = 0.0;
        if (fabs(q) > kZeroPlus) {
            const int bx = sI[0][a] - S + 1 + sp.K[0], by = sI[1][a] - S + 1 + sp.K[1], bz = sI[2][a] - S + 1 + sp.K[2];
            for (int p = lane; p < S * S; p += 64) {
                const int tx = p % S, ty = p / S;
                const int gx = (bx + tx) % sp.K[0], gy = (by + ty) % sp.K[1];
                const R mx = sM[0][a][tx], my = sM[1][a][ty], dx = sD[0][a][tx], dy = sD[1][a][ty];
                R se = (R)0, sx = (R)0, sy = (R)0, sz = (R)0;
#pragma unroll
                for (int tz = 0; tz < S; tz++) {
                    const int gz = (bz + tz) % sp.K[2];
                    const R phi = Phi[((size_t)gz * sp.K[1] + gy) * sp.K[0] + gx];
                    const R mz = sM[2][a][tz], dz = sD[2][a][tz];
                    se += mz * phi;
                    sz += dz * phi;
                }
                // separable in z: sum_z M_z phi and sum_z dM_z phi
                sx = dx * my * se;
                sy = mx * dy * se;
                e += (double)(mx * my * se);
                fx += (double)sx;
                fy += (double)sy;
                fz += (double)(mx * my * sz);
            }
        }
        e = wave_sum(e);
        fx = wave_sum(fx);
        fy = wave_sum(fy);
        fz = wave_sum(fz);
        if (lane == 0) {
            // ewald_general.F90:800-855: F_i = -q sum dM.M.M.Phi * recip_kmax
            const double Fx = -q * fx * sp.rk[0], Fy = -q * fy * sp.rk[1], Fz = -q * fz * sp.rk[2];
            frc_recip[a0 + a] = make_double4(Fx, Fy, Fz, 0.0);
            acc[0] += q * e;
            acc[1] += Fx;
            acc[2] += Fy;
            acc[3] += Fz;
        }
    }
    __syncthreads();
    block_sum<4>(acc, lred, red + (size_t)blockIdx.x * kReduceSlots + R_ERECIP);
}

// total force += 2 scale (F_recip - mean F_recip) (ewald_general.F90:857-864,
// ewald_spole.F90:430)
__global__ void k_finish(double4 *__restrict__ frc, const double4 *__restrict__ frc_recip, int n, double two_scale,
                         const double *__restrict__ red_out, double inv_n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double mx = red_out[R_FNETX] * inv_n, my = red_out[R_FNETY] * inv_n, mz = red_out[R_FNETZ] * inv_n;
    const double4 r = frc_recip[i];
    double4 f = frc[i];
    f.x += (r.x - mx) * two_scale;
    f.y += (r.y - my) * two_scale;
    f.z += (r.z - mz) * two_scale;
    frc[i] = f;
}

// Sum the partial rows of every column into red_out.
struct RedRows {
    int rows[kReduceSlots];
};
// one workgroup per column; fixed summation order (deterministic)
__global__ void __launch_bounds__(256) k_reduce(const double *__restrict__ red, RedRows rr, double *out)
{
    __shared__ double part[256];
    const int col = blockIdx.x;
    double s = 0.0;
    for (int r = threadIdx.x; r < rr.rows[col]; r += blockDim.x) s += red[(size_t)r * kReduceSlots + col];
    part[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) part[threadIdx.x] += part[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[col] = part[0];
}

int ensure_red_rows(Ctx *c, int rows)
{
    if (rows <= c->red_cap) return 0;
    if (c->red) (void)hipFree(c->red);
    c->red = nullptr;
    DLPX_HIP(c, hipMalloc((void **)&c->red, sizeof(double) * kReduceSlots * (size_t)rows));
    c->red_cap = rows;
    return 0;
}

int reduce_partials(Ctx *c)
{
    RedRows rr;
    for (int k = 0; k < kReduceSlots; k++) rr.rows[k] = c->red_rows[k];
    hipLaunchKernelGGL(k_reduce, dim3(kReduceSlots), dim3(256), 0, c->stream, c->red, rr, c->red_out);
    DLPX_HIP(c, hipGetLastError());
    return 0;
}

static SpmeParams spme_params(const Ctx *c)
{
    SpmeParams sp;
    for (int d = 0; d < 3; d++) {
        sp.K[d] = c->K[d];
        sp.kr[d] = (double)c->K[d];
    }
    sp.K1h = c->K1h;
    sp.S = c->prm.nsplines;
    for (int k = 0; k < 9; k++) sp.rcell[k] = c->rcell[k];
    for (int d = 0; d < 3; d++)
        sp.rk[d] = c->rcell[d] * sp.kr[0] + c->rcell[d + 3] * sp.kr[1] + c->rcell[d + 6] * sp.kr[2];
    return sp;
}

int spme_setup(Ctx *c)
{
    const int S = c->prm.nsplines;
    if (S < 4 || S > kMaxSplines || (S & 1))
        return fail(c, DLPX_ERR_UNSUPPORTED, "spline order must be even and in 4..16");
    for (int d = 0; d < 3; d++) {
        int k = c->prm.kvec[d];
        if (k < S) return fail(c, DLPX_ERR_BAD_ARG, "kvec must be >= the spline order");
        for (int p : {2, 3, 5})
            while (k % p == 0) k /= p;
        if (k != 1) return fail(c, DLPX_ERR_BAD_ARG, "kvec entries must be 2^a 3^b 5^c");
        c->K[d] = c->prm.kvec[d];
    }
    c->K1h = c->K[0] / 2 + 1;
    c->nreal = (size_t)c->K[0] * c->K[1] * c->K[2];
    c->nhalf = (size_t)c->K1h * c->K[1] * c->K[2];
    const size_t rs = c->prm.precision == DLPX_MIXED ? sizeof(float) : sizeof(double);
    DLPX_HIP(c, hipMalloc(&c->Q, rs * c->nreal));
    DLPX_HIP(c, hipMalloc(&c->H, 2 * rs * c->nhalf));
    const int ksum = c->K[0] + c->K[1] + c->K[2];
    std::vector<double> n2(ksum);
    bspline_norm2(c->K[0], S, n2.data());
    bspline_norm2(c->K[1], S, n2.data() + c->K[0]);
    bspline_norm2(c->K[2], S, n2.data() + c->K[0] + c->K[1]);
    DLPX_HIP(c, hipMalloc((void **)&c->bsp_n2, sizeof(double) * ksum));
    DLPX_HIP(c, hipMemcpy(c->bsp_n2, n2.data(), sizeof(double) * ksum, hipMemcpyHostToDevice));
    // twiddles exp(+2 pi i m/K) per axis, in the grid precision
    std::vector<double> twd(2 * (size_t)ksum);
    size_t o = 0;
    for (int d = 0; d < 3; d++)
        for (int m = 0; m < c->K[d]; m++, o++) {
            const double a = 2.0 * kPi * (double)m / (double)c->K[d];
            twd[2 * o] = std::cos(a);
            twd[2 * o + 1] = std::sin(a);
        }
    if (c->prm.precision == DLPX_MIXED) {
        std::vector<float> twf(twd.begin(), twd.end());
        DLPX_HIP(c, hipMalloc(&c->tw, sizeof(float) * twf.size()));
        DLPX_HIP(c, hipMemcpy(c->tw, twf.data(), sizeof(float) * twf.size(), hipMemcpyHostToDevice));
    } else {
        DLPX_HIP(c, hipMalloc(&c->tw, sizeof(double) * twd.size()));
        DLPX_HIP(c, hipMemcpy(c->tw, twd.data(), sizeof(double) * twd.size(), hipMemcpyHostToDevice));
    }
    return 0;
}

template <typename R>
static int spme_run(Ctx *c)
{
    const int n = c->natms;
    hipStream_t s = c->stream;
    const SpmeParams sp = spme_params(c);
    const int nb = ceil_div(n, kAtomsPerBlock);
    int r;
    if ((r = ensure_red_rows(c, nb))) return r;
    // ewald_general.F90:517-576 charge spreading
    DLPX_HIP(c, hipMemsetAsync(c->Q, 0, sizeof(R) * c->nreal, s));
    switch (sp.S) {
#define DLPX_SPREAD(SS)                                                                                    \
    case SS:                                                                                               \
        hipLaunchKernelGGL((k_spread_cells<R, SS>), dim3(c->ncell), dim3(256), 0, s, c->pos, c->cell_start, sp, \
                           (R *)c->Q);                                                                     \
        break;
        DLPX_SPREAD(4) DLPX_SPREAD(6) DLPX_SPREAD(8) DLPX_SPREAD(10) DLPX_SPREAD(12) DLPX_SPREAD(14) DLPX_SPREAD(16)
#undef DLPX_SPREAD
        default: return fail(c, DLPX_ERR_UNSUPPORTED, "spline order must be even and in 4..16");
    }
    if (c->timing) DLPX_HIP(c, hipEventRecord(c->ev[2], s));
    if ((r = fft_forward(c))) return r;
    // ewald_spole.F90:1296-1306 cutoff and convergence factors
    InflParams ip;
    for (int d = 0; d < 3; d++) ip.K[d] = c->K[d];
    ip.K1h = c->K1h;
    for (int k = 0; k < 9; k++) ip.rcell[k] = c->rcell[k];
    double rcp[10];
    dcell(c->rcell, rcp);
    ip.rcf = kPi / c->prm.alpha;
    ip.test_fac = std::pow(1.0e-6 / ip.rcf, 2);
    const double m = std::min(std::min((double)c->K[0] * rcp[6], (double)c->K[1] * rcp[7]), (double)c->K[2] * rcp[8]);
    const double cut_off = 0.5 * 1.05 * m;
    ip.cut2 = cut_off * cut_off;
    const int nbi = std::min(ceil_div((long)c->nhalf, 256), 2048);
    if ((r = ensure_red_rows(c, nbi))) return r;
    ip.nblocks = nbi;
    using R2 = typename std::conditional<sizeof(R) == 8, double2, float2>::type;
    hipLaunchKernelGGL(k_influence<R>, dim3(nbi), dim3(256), 0, s, (R2 *)c->H, c->bsp_n2, ip, c->red);
    for (int k = R_RS11; k <= R_RS33; k++) c->red_rows[k] = nbi;
    if ((r = fft_inverse(c))) return r;
    if (c->timing) DLPX_HIP(c, hipEventRecord(c->ev[3], s));
    switch (sp.S) {
#define DLPX_GATHER(SS)                                                                                     \
    case SS:                                                                                                \
        hipLaunchKernelGGL((k_gather<R, SS>), dim3(nb), dim3(256), 0, s, c->pos, n, sp, (const R *)c->Q,   \
                           c->frc_recip, c->red);                                                           \
        break;
        DLPX_GATHER(4) DLPX_GATHER(6) DLPX_GATHER(8) DLPX_GATHER(10) DLPX_GATHER(12) DLPX_GATHER(14) DLPX_GATHER(16)
#undef DLPX_GATHER
        default: return fail(c, DLPX_ERR_UNSUPPORTED, "spline order must be even and in 4..16");
    }
    for (int k = R_ERECIP; k <= R_FNETZ; k++) c->red_rows[k] = nb;
    if (c->timing) DLPX_HIP(c, hipEventRecord(c->ev[4], s));
    DLPX_HIP(c, hipGetLastError());
    return 0;
}

int spme_forces(Ctx *c)
{
    if (c->prm.precision == DLPX_MIXED) return spme_run<float>(c);
    return spme_run<double>(c);
}

int spme_finish(Ctx *c, double two_scale)
{
    hipLaunchKernelGGL(k_finish, dim3(ceil_div(c->natms, 256)), dim3(256), 0, c->stream, c->frc, c->frc_recip,
                       c->natms, two_scale, c->red_out, 1.0 / (double)c->natms);
    DLPX_HIP(c, hipGetLastError());
    return 0;
}

}  // namespace dlpx
