// SPME reciprocal space (ewald_spole.F90:244-477, ewald_general.F90:517-869,
// bspline.F90:192-306, ewald_spole.F90:1257-1386) on the GPU.
//
//  k_spread_cells  one workgroup per link cell: B-splines of its atoms go to
//               LDS and each wavefront adds one atom's S^3 charge weights into
//               an LDS window of the grid (lanes over the (x, y) plane), which
//               is then added to the global grid.
//  (fft.hip)    Q <- FFT-(theta . FFT+(Q)); theta(m) on the half spectrum
//               with the spherical cutoff 0.525 min_d K_d w*_d, and the
//               reciprocal stress.
//  k_gather     B-splines of a block's atoms staged in LDS; a group of lanes
//               per atom (lane = x column) reads the potential grid, a group
//               reduction gives the atom's energy and force.
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
    float invK[3];          // 1/K (fast wrapping of grid indices)
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
template <typename R, bool DERIV, int S, bool DERIV2 = false>
__device__ __forceinline__ void bspline(R w, R *M, R *D, R *D2 = nullptr)
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
    // l = 2 stage: second derivatives from order S-2 (z_j - 2 z_j+1 + z_j+2,
    // zd[S+1] = 0), then raise to order S-1
    {
        const R k_r = (R)(S - 1);
        const R km1_rr = (R)1 / (R)(S - 2);
#pragma unroll
        for (int j = 1; j <= S - 1; j++) {
            if (DERIV2) D2[j - 1] = (zd[j] - (R)2 * zd[j + 1]) + zd[j + 2];
            const R pt = w + (R)(S - j);
            zd[j] = (pt * zd[j] + (k_r - pt) * zd[j + 1]) * km1_rr;
        }
        if (DERIV2) D2[S - 1] = zd[S];
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

// Unwrapped fractional grid coordinate of one axis (no folding): the base
// grid point floor(u) - S + 1 may be negative or >= K; callers fold indices.
__device__ __forceinline__ double grid_u(const SpmeParams &sp, int d, double x, double y, double z)
{
    return sp.kr[d] * (sp.rcell[d] * x + sp.rcell[d + 3] * y + sp.rcell[d + 6] * z + 0.5);
}

// floor(u) for the spreading window, held within +-2^20: a coordinate far
// outside the cell (a diverging run) then still gives window bounds whose
// integer arithmetic cannot overflow (an overflowed bound passed the window's
// one-period test and indexed the grid out of range); such a window exceeds
// the mesh and takes the direct path, which wraps every index.  NaN maps to
// the lower bound (fmax returns the non-NaN operand).
__device__ __forceinline__ double spread_floor(double u)
{
    return fmin(fmax(floor(u), -1048576.0), 1048576.0);
}

// A value every lane of the wavefront holds, moved to scalar registers.
__device__ __forceinline__ float uniform_value(float v)
{
    return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(v)));
}
__device__ __forceinline__ double uniform_value(double v)
{
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readfirstlane((int)(b & 0xffffffffll)), hi = __builtin_amdgcn_readfirstlane((int)(b >> 32));
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
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
//
// The window accumulates in FP64 at either precision: FP64 LDS atomic adds
// (ds_add_f64) sustain several times the rate of FP32 ones (ds_add_f32) on
// gfx950, and the FP32 grid then receives one rounding per window point.
// One workgroup per link cell gives only ~730 workgroups at 64k ions, so each
// needs many wavefronts to keep the LDS atomics busy: 512 threads in FP32;
// the FP64 build's register use makes 256 the better trade there (measured).
//
// The B-splines of up to spread_atoms() atoms are staged in LDS once and
// serve every z-slab; a cell with more atoms restages them per slab.  Each
// wavefront adds one atom at a time, lanes over the (x, y) points, with the
// atom's S z-weights held in registers so the S atomic adds of a lane issue
// back to back.
template <typename R>
constexpr int spread_threads()
{
    return sizeof(R) == 4 ? 512 : 256;
}
template <typename R, int S>
constexpr int spread_atoms()
{
    // 12 KB (FP32) or 18 KB (FP64) of staged splines: with the 32 KB slab,
    // three workgroups fit in a CU's LDS
    return (sizeof(R) == 8 ? 18432 : 12288) / (3 * S * (int)sizeof(R));
}
template <typename R>
struct SpreadArgs {
    const double4 *pos;
    const int32_t *cell_start;
    int ncell;
    SpmeParams sp;
    R *Q;
};
template <typename R, int S>
__device__ __forceinline__ void spread_body(const SpreadArgs<R> &sa, int c)
{
    constexpr int kCap = kSlabBytes / (int)sizeof(double);
    constexpr int NA = spread_atoms<R, S>();
    constexpr int NT = spread_threads<R>();
    __shared__ double slab[kCap];
    __shared__ R sM[3][NA][S];
    __shared__ int sB[3][NA];
    __shared__ R sQ[NA];
    __shared__ int wlo[3], whi[3];
    const double4 *__restrict__ pos = sa.pos;
    const int32_t *__restrict__ cell_start = sa.cell_start;
    const SpmeParams &sp = sa.sp;
    R *__restrict__ Q = sa.Q;
    const int t = threadIdx.x, wid = t >> 6, lane = t & 63;
    if (c >= sa.ncell) return;
    const int s0 = cell_start[c], n = cell_start[c + 1] - s0;
    if (n <= 0) return;
    if (t < 3) {
        wlo[t] = 0x7fffffff;
        whi[t] = -0x7fffffff;
    }
    __syncthreads();
    // window bounds; a cell whose atoms all fit the staging arrays has their
    // B-splines, base points and charges staged in the same pass (one read of
    // the positions)
    const bool once = n <= NA;
    // per-thread bounds, reduced across the wavefront before one LDS atomic
    // per wavefront: atomics of every atom on the same six LDS words
    // serialise (measured: 9 us of the kernel at 64k ions)
    int bmin[3] = {0x7fffffff, 0x7fffffff, 0x7fffffff}, bmax[3] = {-0x7fffffff, -0x7fffffff, -0x7fffffff};
    for (int k = t; k < n; k += NT) {
        const double4 p = pos[s0 + k];
        for (int d = 0; d < 3; d++) {
            const double u = grid_u(sp, d, p.x, p.y, p.z);
            const double fl = spread_floor(u);
            const int b = (int)fl - S + 1;
            bmin[d] = min(bmin[d], b);
            bmax[d] = max(bmax[d], b);
            if (once) {
                R M[S];
                bspline<R, false, S>((R)(u - fl), M, nullptr);
#pragma unroll
                for (int m = 0; m < S; m++) sM[d][k][m] = M[m];
                sB[d][k] = b;
            }
        }
        if (once) sQ[k] = (R)p.w;
    }
#pragma unroll
    for (int d = 0; d < 3; d++) {
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
            bmin[d] = min(bmin[d], __shfl_xor(bmin[d], o, 64));
            bmax[d] = max(bmax[d], __shfl_xor(bmax[d], o, 64));
        }
        if (lane == 0 && bmin[d] <= bmax[d]) {
            atomicMin(&wlo[d], bmin[d]);
            atomicMax(&whi[d], bmax[d]);
        }
    }
    __syncthreads();
    const int lo0 = wlo[0], lo1 = wlo[1], lo2 = wlo[2];
    // window extents in 64-bit so that non-finite coordinates cannot overflow
    const long ex = (long)whi[0] - lo0 + S, ey = (long)whi[1] - lo1 + S, ez = (long)whi[2] - lo2 + S;
    const bool tile = ex <= sp.K[0] && ey <= sp.K[1] && ez <= sp.K[2] && ex * ey <= kCap;
    const int Wx = tile ? (int)ex : 0, Wy = tile ? (int)ey : 0, Wz = tile ? (int)ez : 1;
    const int Dz = tile ? min(Wz, kCap / (Wx * Wy)) : Wz;
    const int plane = Wx * Wy;
    // B-splines, base points and charges of atoms a0 .. a0+NA-1 into LDS
    auto stage = [&](int a0) {
        for (int k = t; k < 3 * NA; k += NT) {
            const int d = k / NA, a = k % NA;
            if (a0 + a < n) {
                const double4 p = pos[s0 + a0 + a];
                const double u = grid_u(sp, d, p.x, p.y, p.z);
                const double fl = spread_floor(u);
                R M[S];
                bspline<R, false, S>((R)(u - fl), M, nullptr);
#pragma unroll
                for (int m = 0; m < S; m++) sM[d][a][m] = M[m];
                sB[d][a] = (int)fl - S + 1;
                if (d == 0) sQ[a] = (R)p.w;
            }
        }
    };
    for (int z0 = 0; z0 < Wz; z0 += Dz) {
        if (tile) {
            if (z0 > 0) __syncthreads();  // the previous slab's flush has read the slab
            for (int k = t; k < plane * Dz; k += NT) slab[k] = 0.0;
        }
        for (int a0 = 0; a0 < n; a0 += NA) {
            __syncthreads();
            if (!once) {
                stage(a0);
                __syncthreads();
            }
            const int na = min(NA, n - a0);
            for (int a = wid; a < na; a += NT / 64) {
                // the atom's charge as a wavefront-uniform value: the skip below
                // is then a scalar branch
                const R q = uniform_value(sQ[a]);
                if (fabs((double)q) <= kZeroPlus) continue;
                // one atom per wavefront: its base point and z-range are scalars,
                // so the per-z tests below are scalar branches, not exec masks
                const int bx = __builtin_amdgcn_readfirstlane(sB[0][a]);
                const int by = __builtin_amdgcn_readfirstlane(sB[1][a]);
                const int bz = __builtin_amdgcn_readfirstlane(sB[2][a]);
                // spline points of this atom inside the current z-slab
                const int tz0 = tile ? max(0, z0 + lo2 - bz) : 0;
                const int tz1 = tile ? min(S, z0 + lo2 + Dz - bz) : S;
                if (tz0 >= tz1) continue;
                R mz[S];
#pragma unroll
                for (int m = 0; m < S; m++) mz[m] = q * sM[2][a][m];
                for (int pp = lane; pp < S * S; pp += 64) {
                    const int tx = pp % S, ty = pp / S;
                    const R mxy = sM[0][a][tx] * sM[1][a][ty];
                    if (tile) {
                        // slab index of spline point tz = 0 (points outside the slab are skipped);
                        // the slab's spline points as a scalar bit mask, one bit test per point
                        const int l0 = (bz - lo2 - z0) * plane + (by + ty - lo1) * Wx + (bx + tx - lo0);
                        const unsigned zm = ((1u << tz1) - 1u) & ~((1u << tz0) - 1u);
#pragma unroll
                        for (int tz = 0; tz < S; tz++)
                            if (zm & (1u << tz)) atomicAdd(&slab[l0 + tz * plane], (double)(mz[tz] * mxy));
                    } else {
                        const int gx = wrapk(bx + tx, sp.K[0]), gy = wrapk(by + ty, sp.K[1]);
#pragma unroll
                        for (int tz = 0; tz < S; tz++) {
                            const int gz = wrapk(bz + tz, sp.K[2]);
                            unsafeAtomicAdd(&Q[((size_t)gz * sp.K[1] + gy) * sp.K[0] + gx], mz[tz] * mxy);
                        }
                    }
                }
            }
        }
        __syncthreads();
        if (!tile) break;  // the direct path did every z at once
        const int dz = min(Dz, Wz - z0);
        // window coordinates within one period of the grid (always, for
        // positions near the cell) wrap with one compare
        const bool near = lo0 >= -sp.K[0] && lo1 >= -sp.K[1] && lo2 >= -sp.K[2] && lo0 + Wx <= 2 * sp.K[0] &&
                          lo1 + Wy <= 2 * sp.K[1] && lo2 + Wz <= 2 * sp.K[2];
        if (near && Wx <= 64) {
            // one (y, z) row of the slab per group of 32 or 64 lanes, lanes over x
            const int LW = Wx <= 32 ? 32 : 64, lx = lane & (LW - 1);
            auto fold = [](int v, int K) { return v < 0 ? v + K : (v >= K ? v - K : v); };
            const int gx = fold(lo0 + lx, sp.K[0]);
            for (int row = wid * (64 / LW) + lane / LW; row < Wy * dz; row += NT / LW) {
                const int lz = row / Wy, ly = row - lz * Wy;
                const R v = lx < Wx ? (R)slab[row * Wx + lx] : (R)0;
                if (v == (R)0) continue;
                const int gy = fold(lo1 + ly, sp.K[1]), gz = fold(lo2 + z0 + lz, sp.K[2]);
                unsafeAtomicAdd(&Q[((size_t)gz * sp.K[1] + gy) * sp.K[0] + gx], v);
            }
            continue;
        }
        const float iWx = 1.0f / (float)Wx, iWy = 1.0f / (float)Wy;
        // the float-reciprocal wrap is exact for window coordinates below 2^20
        // in magnitude
        const bool small = abs(lo0) < (1 << 19) && abs(lo1) < (1 << 19) && abs(lo2) < (1 << 19);
        for (int k = t; k < plane * dz; k += NT) {
            const R v = (R)slab[k];
            if (v == (R)0) continue;
            int lx, ly, gx, gy, gz;
            const int lz = fdiv_floor(fdiv_floor(k, Wx, iWx, lx), Wy, iWy, ly);
            if (small) {
                (void)fdiv_floor(lo0 + lx, sp.K[0], sp.invK[0], gx);
                (void)fdiv_floor(lo1 + ly, sp.K[1], sp.invK[1], gy);
                (void)fdiv_floor(lo2 + z0 + lz, sp.K[2], sp.invK[2], gz);
            } else {
                gx = wrapk(lo0 + lx, sp.K[0]);
                gy = wrapk(lo1 + ly, sp.K[1]);
                gz = wrapk(lo2 + z0 + lz, sp.K[2]);
            }
            unsafeAtomicAdd(&Q[((size_t)gz * sp.K[1] + gy) * sp.K[0] + gx], v);
        }
    }
}

template <typename R, int S>
__global__ void __launch_bounds__(spread_threads<R>()) k_spread_cells(SpreadArgs<R> a)
{
    spread_body<R, S>(a, xcd_block(blockIdx.x, gridDim.x));
}
template <typename R, int S>
__global__ void __launch_bounds__(spread_threads<R>()) k_spread_cells_batch(const SpreadArgs<R> *__restrict__ av,
                                                                            const int32_t *__restrict__ boff,
                                                                            int nsys)
{
    BatchPos bp;
    if (!batch_pos(boff, nsys, bp)) return;
    const SpreadArgs<R> &a = av[bp.s];
    spread_body<R, S>(a, bp.b);
}

// Lanes per atom in the gather: the next power of two >= S, lane tx < S
// taking the grid column x = base + tx.
constexpr int gather_lanes(int S) { return S <= 4 ? 4 : S <= 8 ? 8 : 16; }

// Force interpolation (ewald_general.F90:776-855).  A group of gather_lanes(S)
// lanes serves one atom: lane tx reads the S x S (y, z) potential values of
// its grid column, so one wavefront load instruction covers 64/G atoms with
// runs of S contiguous x points each, and the cross-lane reduction spans only
// the group.  z and y sums are separable:
//   E  = sum_y M_y sum_z M_z Phi,  Fy = sum_y dM_y sum_z M_z Phi,
//   Fz = sum_y M_y sum_z dM_z Phi,
// then e = M_x E, fx = dM_x E, fy = M_x Fy, fz = M_x Fz per lane.  The z sums
// run in R, everything after them in FP64.
template <typename R>
struct GatherArgs {
    const double4 *pos;
    int n;
    SpmeParams sp;
    const R *Phi;
    double4 *frc_recip;
    double *red;
    // batched runs: the energy and net-force totals are also added into
    // acc[4 (bid % kAccSlots) + 0..3] (zero before the launch), one atomic
    // each per workgroup, so the velocity stage needs no reduction launch to
    // subtract the mean reciprocal force (null: rows only)
    double *acc;
};
template <typename R, int S>
__device__ __forceinline__ void gather_body(const GatherArgs<R> &ga, int bid)
{
    constexpr int G = gather_lanes(S), APB = 256 / G;
    __shared__ R sM[3][APB][S];
    __shared__ R sD[3][APB][S];
    __shared__ int sB[3][APB];
    __shared__ double sQ[APB];
    __shared__ double lred[4 * 4];
    const double4 *__restrict__ pos = ga.pos;
    const int n = ga.n;
    const SpmeParams &sp = ga.sp;
    const R *__restrict__ Phi = ga.Phi;
    double4 *__restrict__ frc_recip = ga.frc_recip;
    const int a0 = bid * APB;
    const int t = threadIdx.x;
    if (t < 3 * APB) {
        const int d = t / APB, a = t % APB;
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
            // first grid point of the stencil, in [0, K)
            int b = idx - S + 1;
            if (b < 0) b += sp.K[d];
            sB[d][a] = b;
            if (d == 0) sQ[a] = p.w;
        }
    }
    __syncthreads();
    const int a = t / G, tx = t % G, i = a0 + a;
    double acc[4] = {0, 0, 0, 0};  // energy, fx, fy, fz (unscaled, summed over atoms)
    double e = 0.0, fx = 0.0, fy = 0.0, fz = 0.0;
    const double q = i < n ? sQ[a] : 0.0;
    if (i < n && tx < S && fabs(q) > kZeroPlus) {
        int gx = sB[0][a] + tx;
        if (gx >= sp.K[0]) gx -= sp.K[0];
        const int by = sB[1][a], bz = sB[2][a];
        const size_t plane = (size_t)sp.K[0] * sp.K[1];
        // z planes of the stencil, wrapped once
        size_t zoff[S];
#pragma unroll
        for (int tz = 0; tz < S; tz++) {
            int gz = bz + tz;
            if (gz >= sp.K[2]) gz -= sp.K[2];
            zoff[tz] = (size_t)gz * plane + gx;
        }
        double E = 0.0, Fy = 0.0, Fz = 0.0;
        // y rows per iteration (a divisor of S, which is even), their RY*S
        // grid values in flight together
        constexpr int RY = S % 4 == 0 ? 4 : 2;
        for (int ty = 0; ty < S; ty += RY) {
            size_t yo[RY];
#pragma unroll
            for (int r = 0; r < RY; r++) {
                int gy = by + ty + r;
                if (gy >= sp.K[1]) gy -= sp.K[1];
                yo[r] = (size_t)gy * sp.K[0];
            }
            R ph[RY][S];
#pragma unroll
            for (int tz = 0; tz < S; tz++)
#pragma unroll
                for (int r = 0; r < RY; r++) ph[r][tz] = Phi[zoff[tz] + yo[r]];
#pragma unroll
            for (int r = 0; r < RY; r++) {
                R se = (R)0, sz = (R)0;
#pragma unroll
                for (int tz = 0; tz < S; tz++) {
                    se += sM[2][a][tz] * ph[r][tz];
                    sz += sD[2][a][tz] * ph[r][tz];
                }
                const double my = (double)sM[1][a][ty + r], dy = (double)sD[1][a][ty + r];
                E += my * (double)se;
                Fy += dy * (double)se;
                Fz += my * (double)sz;
            }
        }
        const double mx = (double)sM[0][a][tx], dx = (double)sD[0][a][tx];
        e = mx * E;
        fx = dx * E;
        fy = mx * Fy;
        fz = mx * Fz;
    }
    e = group_sum<G>(e);
    fx = group_sum<G>(fx);
    fy = group_sum<G>(fy);
    fz = group_sum<G>(fz);
    if (i < n && tx == 0) {
        // ewald_general.F90:800-855: F_i = -q sum dM.M.M.Phi * recip_kmax
        const double Fx_ = -q * fx * sp.rk[0], Fy_ = -q * fy * sp.rk[1], Fz_ = -q * fz * sp.rk[2];
        frc_recip[i] = make_double4(Fx_, Fy_, Fz_, q * e);  // w: the atom's energy (ewald_general.F90:852)
        acc[0] = q * e;
        acc[1] = Fx_;
        acc[2] = Fy_;
        acc[3] = Fz_;
    }
    const double tot = block_sum<4>(acc, lred, ga.red + (size_t)bid * kReduceSlots + R_ERECIP);
    if (ga.acc && t < 4 && tot != 0.0) atomicAdd(&ga.acc[4 * (bid % kAccSlots) + t], tot);
}
template <typename R, int S>
__global__ void __launch_bounds__(256) k_gather(GatherArgs<R> a)
{
    gather_body<R, S>(a, xcd_block(blockIdx.x, gridDim.x));
}
template <typename R, int S>
__global__ void __launch_bounds__(256) k_gather_batch(const GatherArgs<R> *__restrict__ av,
                                                      const int32_t *__restrict__ boff, int nsys)
{
    BatchPos bp;
    if (!batch_pos(boff, nsys, bp)) return;
    const GatherArgs<R> &a = av[bp.s];
    gather_body<R, S>(a, bp.b);
}

// Per-particle reciprocal stress (spme_calc_stress, ewald_general.F90:871-
// 1009) from the stress grid G (the charge grid convolved with the stress
// kernel): S_ab = -q sum (d_a d_b M_x M_y M_z) G kmax_a kmax_b / (4 pi alpha)
// with second derivatives of the splines, for ab = 11, 21, 31, 22, 32, 33.
// The lane layout of k_gather: lane tx takes grid column x; the (y, z) sums
//   A = sum My Mz G, B = sum dMy Mz G, C = sum My dMz G,
//   D = sum d2My Mz G, E = sum dMy dMz G, F = sum My d2Mz G
// give S11 ~ d2Mx A, S21 ~ dMx B, S31 ~ dMx C, S22 ~ Mx D, S32 ~ Mx E,
// S33 ~ Mx F.  Writes pp[i*kPPStride + 8..13] and the block sums of the six
// components into R_RS11..R_RS33: with per-particle collection the reference
// takes the reciprocal stress from these sums (ewald_general.F90:963,1002).
template <typename R, int S>
__global__ void __launch_bounds__(256) k_gather_stress(const double4 *__restrict__ pos, int n, SpmeParams sp,
                                                       double c_fac, const R *__restrict__ Gs,
                                                       double *__restrict__ pp, double *red)
{
    constexpr int G = gather_lanes(S), APB = 256 / G;
    __shared__ R sM[3][APB][S];
    __shared__ R sD[3][APB][S];
    __shared__ R sD2[3][APB][S];
    __shared__ int sB[3][APB];
    __shared__ double sQ[APB];
    __shared__ double lred[6 * 4];
    const int a0 = xcd_block(blockIdx.x, gridDim.x) * APB;
    const int t = threadIdx.x;
    if (t < 3 * APB) {
        const int d = t / APB, a = t % APB;
        const int i = a0 + a;
        if (i < n) {
            const double4 p = pos[i];
            int idx;
            double w;
            grid_coord(sp, d, p.x, p.y, p.z, idx, w);
            R M[S], D[S], D2[S];
            bspline<R, true, S, true>((R)w, M, D, D2);
#pragma unroll
            for (int k = 0; k < S; k++) {
                sM[d][a][k] = M[k];
                sD[d][a][k] = D[k];
                sD2[d][a][k] = D2[k];
            }
            int b = idx - S + 1;
            if (b < 0) b += sp.K[d];
            sB[d][a] = b;
            if (d == 0) sQ[a] = p.w;
        }
    }
    __syncthreads();
    const int a = t / G, tx = t % G, i = a0 + a;
    double v[6] = {0, 0, 0, 0, 0, 0};
    if (i < n && tx < S) {
        int gx = sB[0][a] + tx;
        if (gx >= sp.K[0]) gx -= sp.K[0];
        const int by = sB[1][a], bz = sB[2][a];
        const size_t plane = (size_t)sp.K[0] * sp.K[1];
        double A = 0, B = 0, C = 0, D = 0, E = 0, F = 0;
        for (int tz = 0; tz < S; tz++) {
            int gz = bz + tz;
            if (gz >= sp.K[2]) gz -= sp.K[2];
            const double mz = (double)sM[2][a][tz], dz = (double)sD[2][a][tz], d2z = (double)sD2[2][a][tz];
            double a_ = 0, b_ = 0, d_ = 0;  // sum over y of My G, dMy G, d2My G
            for (int ty = 0; ty < S; ty++) {
                int gy = by + ty;
                if (gy >= sp.K[1]) gy -= sp.K[1];
                const double g = (double)Gs[(size_t)gz * plane + (size_t)gy * sp.K[0] + gx];
                a_ += (double)sM[1][a][ty] * g;
                b_ += (double)sD[1][a][ty] * g;
                d_ += (double)sD2[1][a][ty] * g;
            }
            A += mz * a_;
            B += mz * b_;
            C += dz * a_;
            D += mz * d_;
            E += dz * b_;
            F += d2z * a_;
        }
        const double mx = (double)sM[0][a][tx], dx = (double)sD[0][a][tx], d2x = (double)sD2[0][a][tx];
        v[0] = d2x * A;
        v[1] = dx * B;
        v[2] = dx * C;
        v[3] = mx * D;
        v[4] = mx * E;
        v[5] = mx * F;
    }
#pragma unroll
    for (int k = 0; k < 6; k++) v[k] = group_sum<G>(v[k]);
    double acc[6] = {0, 0, 0, 0, 0, 0};
    if (i < n && tx == 0) {
        const double q = sQ[a];
        const int ab[6][2] = {{0, 0}, {1, 0}, {2, 0}, {1, 1}, {2, 1}, {2, 2}};
#pragma unroll
        for (int k = 0; k < 6; k++) {
            acc[k] = -(q * v[k]) * sp.rk[ab[k][0]] * sp.rk[ab[k][1]] * c_fac;
            pp[(size_t)i * kPPStride + 8 + k] = acc[k];
        }
    }
    block_sum<6>(acc, lred, red + (size_t)blockIdx.x * kReduceSlots + R_RS11);
}

int gather_blocks(const Ctx *c)
{
    return xcd_grid(ceil_div(c->natms, 256 / gather_lanes(c->prm.nsplines)));
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
__device__ __forceinline__ void reduce_body(const double *__restrict__ red, const int *rows, double *out, int col,
                                            double *rec)
{
    __shared__ double part[256];
    const double v = column_sum<256>(red, rows[col], kReduceSlots, col, part);
    if (threadIdx.x == 0) {
        out[col] = v;
        if (rec) rec[col] = v;
    }
}
__global__ void __launch_bounds__(256) k_reduce(const double *__restrict__ red, RedRows rr, double *out)
{
    reduce_body(red, rr.rows, out, blockIdx.x, nullptr);
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
        sp.invK[d] = 1.0f / (float)c->K[d];
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
    c->K1hp = (c->K1h + 15) / 16 * 16;  // row stride of the half spectrum (128-byte aligned runs)
    c->nhalf = (size_t)c->K1hp * c->K[1] * c->K[2];
    const size_t rs = c->prm.precision == DLPX_MIXED ? sizeof(float) : sizeof(double);
    DLPX_HIP(c, hipMalloc(&c->Q, rs * c->nreal));
    DLPX_HIP(c, dev_zero(c->stream, c->Q, rs * c->nreal));  // kept zero from here on by the forward x pass
    DLPX_HIP(c, hipMalloc(&c->Phi, rs * c->nreal));
    DLPX_HIP(c, hipMalloc(&c->H, 2 * rs * c->nhalf));
    DLPX_HIP(c, hipMalloc(&c->theta, 2 * rs * c->nhalf));
    DLPX_HIP(c, dev_zero(c->stream, c->theta, 2 * rs * c->nhalf));  // the row padding stays zero
    c->theta_valid = false;
    const int ksum = c->K[0] + c->K[1] + c->K[2];
    std::vector<double> n2(ksum);
    bspline_norm2(c->K[0], S, n2.data());
    bspline_norm2(c->K[1], S, n2.data() + c->K[0]);
    bspline_norm2(c->K[2], S, n2.data() + c->K[0] + c->K[1]);
    DLPX_HIP(c, hipMalloc((void **)&c->bsp_n2, sizeof(double) * ksum));
    DLPX_HIP(c, dev_upload(c->stream, c->bsp_n2, n2.data(), sizeof(double) * ksum));
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
        DLPX_HIP(c, dev_upload(c->stream, c->tw, twf.data(), sizeof(float) * twf.size()));
    } else {
        DLPX_HIP(c, hipMalloc(&c->tw, sizeof(double) * twd.size()));
        DLPX_HIP(c, dev_upload(c->stream, c->tw, twd.data(), sizeof(double) * twd.size()));
    }
    return 0;
}

// ewald_spole.F90:1296-1306 cutoff and convergence factors
InflParams infl_params(const Ctx *c)
{
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
    ip.alpha = c->prm.alpha;
    ip.mode = 0;
    return ip;
}

template <typename R>
static SpreadArgs<R> spread_args(const Ctx *c)
{
    return SpreadArgs<R>{c->pos, c->cell_start, c->ncell, spme_params(c), (R *)c->Q};
}
template <typename R>
static GatherArgs<R> gather_args(const Ctx *c)
{
    return GatherArgs<R>{c->pos, c->natms, spme_params(c), (const R *)c->Phi, c->frc_recip, c->red, nullptr};
}

#define DLPX_S_SWITCH(S, CALL)                                                                             \
    switch (S) {                                                                                           \
        case 4: CALL(4); break;                                                                            \
        case 6: CALL(6); break;                                                                            \
        case 8: CALL(8); break;                                                                            \
        case 10: CALL(10); break;                                                                          \
        case 12: CALL(12); break;                                                                          \
        case 14: CALL(14); break;                                                                          \
        case 16: CALL(16); break;                                                                          \
        default: return fail(c, DLPX_ERR_UNSUPPORTED, "spline order must be even and in 4..16");           \
    }

template <typename R>
static int spme_run(Ctx *c)
{
    const int n = c->natms;
    hipStream_t s = c->sp;  // the evaluation's reciprocal-space stream
    const SpmeParams sp = spme_params(c);
    const int nb = gather_blocks(c);  // evaluate() has sized the partial-sum rows
    int r;
    // ewald_general.F90:517-576 charge spreading into the charge grid, which
    // is zero here: set at allocation, and cleared by the forward x pass of
    // each convolution after reading it
    const SpreadArgs<R> spa = spread_args<R>(c);
#define DLPX_SPREAD(SS) \
    hipLaunchKernelGGL((k_spread_cells<R, SS>), dim3(xcd_grid(c->ncell)), dim3(spread_threads<R>()), 0, s, spa)
    DLPX_S_SWITCH(sp.S, DLPX_SPREAD)
#undef DLPX_SPREAD
    if (c->timing) DLPX_HIP(c, hipEventRecord(c->ev[2], s));
    InflParams ip = infl_params(c);
    if (c->pp_now) {
        // keep the charge grid for the per-particle stress convolution
        const size_t bytes = (c->prm.precision == DLPX_MIXED ? sizeof(float) : sizeof(double)) * c->nreal;
        DLPX_HIP(c, hipMemcpyAsync(c->Q2, c->Q, bytes, hipMemcpyDeviceToDevice, s));
    }
    if ((r = fft_convolve(c, ip, c->Q, c->Phi, true))) return r;
    if (c->timing) DLPX_HIP(c, hipEventRecord(c->ev[3], s));
    const GatherArgs<R> ga = gather_args<R>(c);
#define DLPX_GATHER(SS) hipLaunchKernelGGL((k_gather<R, SS>), dim3(nb), dim3(256), 0, s, ga)
    DLPX_S_SWITCH(sp.S, DLPX_GATHER)
#undef DLPX_GATHER
    for (int k = R_ERECIP; k <= R_FNETZ; k++) c->red_rows[k] = nb;
    if (c->pp_now) {
        // per-particle stresses: the saved charge grid convolved with the
        // stress kernel, then gathered with second spline derivatives
        InflParams ips = ip;
        ips.mode = 1;
        if ((r = fft_convolve(c, ips, c->Q2, c->Q2, false))) return r;
        const double c_fac = 1.0 / (2.0 * 2.0 * kPi * c->prm.alpha);  // ewald_general.F90:932-933
#define DLPX_GSTRESS(SS)                                                                                        \
    hipLaunchKernelGGL((k_gather_stress<R, SS>), dim3(nb), dim3(256), 0, s, c->pos, n, sp, c_fac,              \
                       (const R *)c->Q2, c->pp, c->red)
        DLPX_S_SWITCH(sp.S, DLPX_GSTRESS)
#undef DLPX_GSTRESS
        for (int k = R_RS11; k <= R_RS33; k++) c->red_rows[k] = nb;
    }
    if (c->timing) DLPX_HIP(c, hipEventRecord(c->ev[4], s));
    DLPX_HIP(c, hipGetLastError());
    return 0;
}

int gather_blocks_exact(const Ctx *c) { return ceil_div(c->natms, 256 / gather_lanes(c->prm.nsplines)); }

// Batched reciprocal space over the systems with electrostatics: spreading,
// the convolution passes and the gather, one launch each.  Every system of
// the batch must have the same spline order.
template <typename R>
static int spme_prepare_t(const std::vector<Ctx *> &cs, SpmeBatch &sb)
{
    std::vector<SpreadArgs<R>> spa;
    std::vector<GatherArgs<R>> ga;
    std::vector<InflParams> ips;
    std::vector<int> nbs, nbg;
    int S = 0;
    for (Ctx *c : cs) {
        const bool on = c->prm.electrostatics != 0;
        if (on) {
            if (S && c->prm.nsplines != S)
                return fail(c, DLPX_ERR_UNSUPPORTED, "systems of one batch must share the spline order");
            S = c->prm.nsplines;
        }
        spa.push_back(on ? spread_args<R>(c) : SpreadArgs<R>{});
        ga.push_back(on ? gather_args<R>(c) : GatherArgs<R>{});
        ga.back().acc = on ? c->acc : nullptr;
        ips.push_back(on ? infl_params(c) : InflParams{});
        nbs.push_back(on ? c->ncell : 0);
        nbg.push_back(on ? gather_blocks_exact(c) : 0);
    }
    sb.S = S;
    int r;
    if ((r = batch_upload(cs[0], sb.spread, spa.data(), sizeof(SpreadArgs<R>), nbs, true)) ||
        (r = batch_upload(cs[0], sb.gather, ga.data(), sizeof(GatherArgs<R>), nbg, true)))
        return r;
    sb.gather_rows = nbg;
    return fft_prepare(cs, ips, sb.fft);
}

int spme_prepare(const std::vector<Ctx *> &cs, SpmeBatch &sb)
{
    if (cs[0]->prm.precision == DLPX_MIXED) return spme_prepare_t<float>(cs, sb);
    return spme_prepare_t<double>(cs, sb);
}

template <typename R>
static int spme_launch_t(Ctx *c, hipStream_t s, const SpmeBatch &sb, hipEvent_t *marks)
{
    if (sb.S == 0) return 0;
    const BatchLaunch &bs = sb.spread, &bg = sb.gather;
#define DLPX_SPREADB(SS)                                                                                         \
    hipLaunchKernelGGL((k_spread_cells_batch<R, SS>), dim3(bs.total), dim3(spread_threads<R>()), 0, s,          \
                       (const SpreadArgs<R> *)bs.args, bs.boff, bs.nsys)
    DLPX_S_SWITCH(sb.S, DLPX_SPREADB)
#undef DLPX_SPREADB
    if (marks && marks[0]) DLPX_HIP(c, hipEventRecord(marks[0], s));
    int r;
    if ((r = fft_launch(c, s, sb.fft))) return r;
    if (marks && marks[1]) DLPX_HIP(c, hipEventRecord(marks[1], s));
#define DLPX_GATHERB(SS)                                                                                         \
    hipLaunchKernelGGL((k_gather_batch<R, SS>), dim3(bg.total), dim3(256), 0, s, (const GatherArgs<R> *)bg.args, \
                       bg.boff, bg.nsys)
    DLPX_S_SWITCH(sb.S, DLPX_GATHERB)
#undef DLPX_GATHERB
    if (marks && marks[2]) DLPX_HIP(c, hipEventRecord(marks[2], s));
    DLPX_HIP(c, hipGetLastError());
    return 0;
}

int spme_launch(Ctx *c, hipStream_t s, const SpmeBatch &sb, hipEvent_t *marks)
{
    if (c->prm.precision == DLPX_MIXED) return spme_launch_t<float>(c, s, sb, marks);
    return spme_launch_t<double>(c, s, sb, marks);
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
