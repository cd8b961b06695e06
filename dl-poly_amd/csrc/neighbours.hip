// Link-cell Verlet neighbour list on the GPU (neighbours.F90:356-1414).
//
// Atoms are binned into link cells of edge >= cutoff+padding, counting-sorted
// by cell (deterministic order inside a cell) and every per-atom array is
// permuted into that order, so the pair and SPME kernels read neighbours that
// sit close in HBM.  The list is a full list (each pair stored from both ends)
// split over G lanes per atom: lane g keeps every G-th candidate it accepts,
// element (i, c, g) lives at nbr[(i*cap_lane + c)*G + g] (list_slot), so one
// atom's entries, and one workgroup's atoms, are contiguous.
#include <cmath>
#include <cstring>

#include "common.h"
#include "device_util.h"

namespace dlpx {

struct CellGrid {
    int nc[3];
    double L[3], invL[3];
};

__device__ __forceinline__ int cell_coord(double x, double L, double invL, int nc)
{
    // x is already folded into [-L/2, L/2)
    int k = (int)((x * invL + 0.5) * (double)nc);
    return k < 0 ? 0 : (k >= nc ? nc - 1 : k);
}

// Runs of equal keys among consecutive active lanes of a wavefront: the
// lane that starts this lane's run and the run's length (all lanes of the
// run get the same answer).  Lanes past the end of the data must have exited.
__device__ __forceinline__ void run_of(int key, int &head, int &len)
{
    const int lane = threadIdx.x & 63;
    const unsigned long long active = __ballot(1);
    const int prev = __shfl_up(key, 1, 64);
    const bool is_head = lane == 0 || prev != key || !((active >> (lane - 1)) & 1ull);
    const unsigned long long heads = __ballot(is_head);
    const unsigned long long upto = (lane == 63) ? ~0ull : ((2ull << lane) - 1ull);
    head = 63 - __clzll(heads & upto);
    const unsigned long long after = heads & ~upto;
    const int next = after ? (__ffsll((long long)after) - 1) : 64;
    // active lanes are a prefix of the wavefront (only the tail wave is partial)
    const int nact = 64 - __clzll(active);
    len = min(next, nact) - head;
}

// Fold positions into the primary cell (relocate_particles / pbcshift) and
// count atoms per link cell.
__global__ void k_wrap_bin(double4 *pos, int n, CellGrid g, int32_t *cell_of, int32_t *count)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double4 p = pos[i];
    p.x -= g.L[0] * floor(p.x * g.invL[0] + 0.5);
    p.y -= g.L[1] * floor(p.y * g.invL[1] + 0.5);
    p.z -= g.L[2] * floor(p.z * g.invL[2] + 0.5);
    // guard the rounding edge x == L/2 after folding
    if (p.x >= 0.5 * g.L[0]) p.x -= g.L[0];
    if (p.y >= 0.5 * g.L[1]) p.y -= g.L[1];
    if (p.z >= 0.5 * g.L[2]) p.z -= g.L[2];
    pos[i] = p;
    const int cx = cell_coord(p.x, g.L[0], g.invL[0], g.nc[0]);
    const int cy = cell_coord(p.y, g.L[1], g.invL[1], g.nc[1]);
    const int cz = cell_coord(p.z, g.L[2], g.invL[2], g.nc[2]);
    const int c = (cz * g.nc[1] + cy) * g.nc[0] + cx;
    cell_of[i] = c;
    // atoms are mostly in cell order already: one atomic per run of equal
    // cells within the wavefront
    int head, len;
    run_of(c, head, len);
    if ((int)(threadIdx.x & 63) == head) atomicAdd(&count[c], len);
}

// Exclusive scan of the cell counts in one workgroup of 1024 threads.
__global__ void __launch_bounds__(1024) k_scan(const int32_t *count, int32_t *start, int32_t *fill, int n)
{
    __shared__ int32_t part[1024];
    __shared__ int32_t carry;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    for (int base = 0; base < n; base += 1024) {
        const int i = base + threadIdx.x;
        const int v = i < n ? count[i] : 0;
        part[threadIdx.x] = v;
        __syncthreads();
        for (int o = 1; o < 1024; o <<= 1) {
            int t = threadIdx.x >= o ? part[threadIdx.x - o] : 0;
            __syncthreads();
            part[threadIdx.x] += t;
            __syncthreads();
        }
        const int incl = part[threadIdx.x];
        if (i < n) {
            start[i] = carry + incl - v;
            fill[i] = carry + incl - v;
        }
        __syncthreads();
        if (threadIdx.x == 1023) carry += incl;
        __syncthreads();
    }
    if (threadIdx.x == 0) start[n] = carry;
}

__global__ void k_scatter(const int32_t *cell_of, int32_t *fill, int32_t *perm, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int c = cell_of[i];
    int head, len;
    run_of(c, head, len);
    const int lane = threadIdx.x & 63;
    int base = 0;
    if (lane == head) base = atomicAdd(&fill[c], len);
    base = __shfl(base, head, 64);
    perm[base + (lane - head)] = i;
}

// Make the order inside each cell deterministic (ascending previous index):
// one workgroup per cell, every member's rank is the number of smaller
// members (indices are distinct), so the sort is one parallel pass in LDS.
constexpr int kSortLds = 2048;
__global__ void __launch_bounds__(256) k_sort_cells(int32_t *perm, const int32_t *start, int ncell)
{
    __shared__ int32_t v[kSortLds];
    const int c = blockIdx.x;
    const int s = start[c], n = start[c + 1] - s;
    if (n <= 1) return;
    if (n > kSortLds) {  // a very crowded cell: serial insertion sort
        if (threadIdx.x == 0)
            for (int a = s + 1; a < s + n; a++) {
                const int x = perm[a];
                int b = a - 1;
                while (b >= s && perm[b] > x) {
                    perm[b + 1] = perm[b];
                    b--;
                }
                perm[b + 1] = x;
            }
        return;
    }
    for (int k = threadIdx.x; k < n; k += blockDim.x) v[k] = perm[s + k];
    __syncthreads();
    for (int k = threadIdx.x; k < n; k += blockDim.x) {
        const int x = v[k];
        int rank = 0;
        for (int j = 0; j < n; j++) rank += v[j] < x;
        perm[s + rank] = x;
    }
}

__global__ void k_permute(const int32_t *perm, int n, const double4 *pos, double4 *pos_new,
                          const int32_t *typ, int32_t *typ_new, const int32_t *gid, int32_t *gid_new,
                          const double4 *vel, double4 *vel_new, int32_t *slot, double4 *xref)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int o = perm[i];
    const double4 p = pos[o];
    pos_new[i] = p;
    xref[i] = p;
    typ_new[i] = typ[o];
    const int g = gid[o];
    gid_new[i] = g;
    slot[g] = i;
    if (vel) vel_new[i] = vel[o];
}

// The permuted arrays back into the primary ones, so device pointers stay
// the same across rebuilds (captured step graphs keep them as arguments).
__global__ void k_copy_back(int n, const double4 *pos_alt, double4 *pos, const int32_t *typ_alt, int32_t *typ,
                            const int32_t *gid_alt, int32_t *gid, const double4 *vel_alt, double4 *vel)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    pos[i] = pos_alt[i];
    typ[i] = typ_alt[i];
    gid[i] = gid_alt[i];
    if (vel) vel[i] = vel_alt[i];
}

// One workgroup per link cell.  The atoms of the (up to) 27 stencil cells
// are staged once in LDS as float4 {x + sx, y + sy, z + sz, j}, with the
// cell's periodic image shift s L already added, and then every atom of the
// cell (G lanes each, lane g taking candidates k = g, g + G, ...) tests them.
//
// Positions are folded into [-L/2, L/2) at the rebuild, so a stencil cell's
// image shift is -L, 0 or +L per axis.  The separation is x_i - (x_j + s L)
// and the criterion r^2 <= (cutoff+padding)^2 (neighbours.F90:991-992).  A
// single-precision distance settles every pair that is clearly inside or
// outside; pairs within a relative 2e-5 of the cutoff are decided in double
// precision without fused multiply-add from the FP64 positions, with the
// arithmetic of the host, so the accepted set is reproducible bit for bit.
// With fewer than 3 cells on an axis the stencil shift need not be the
// minimum image, and every candidate is decided exactly by minimum image.
//
// Stencils with more than kListStage candidates are processed in rounds.
constexpr int kListThreads = 512;
constexpr int kListStage = 3072;  // 48 KB of float4
template <int G>
__global__ void __launch_bounds__(kListThreads) k_build_list(const double4 *__restrict__ pos, int cap, CellGrid g,
                                                             const int32_t *__restrict__ start, double rc2,
                                                             int cap_lane, int32_t *__restrict__ nbr,
                                                             int32_t *__restrict__ cnt, int32_t *flags)
{
    static_assert(G == kLanesPerAtom, "list_slot assumes kLanesPerAtom lanes per atom");
    __shared__ float4 cand[kListStage];
    __shared__ int s_start[27], s_off[28];
    __shared__ double s_sh[27][3];
    __shared__ int s_ovf, s_big, s_tot;  // workgroup partials of flags[0], [1], [3]
    const int c = xcd_block(blockIdx.x, gridDim.x);
    if (c >= g.nc[0] * g.nc[1] * g.nc[2]) return;
    const int t = threadIdx.x;
    if (t == 0) {
        s_ovf = 0;
        s_big = 0;
        s_tot = 0;
    }
    const int cx = c % g.nc[0], cy = (c / g.nc[0]) % g.nc[1], cz = c / (g.nc[0] * g.nc[1]);
    const bool small = g.nc[0] < 3 || g.nc[1] < 3 || g.nc[2] < 3;
    if (t < 27) {
        // stencil cell t: offsets -1..1 per axis, fewer when an axis has 1-2 cells
        const int o[3] = {t % 3 - 1, (t / 3) % 3 - 1, t / 9 - 1};
        const int cc[3] = {cx, cy, cz};
        bool valid = true;
        int jj[3];
        for (int d = 0; d < 3; d++) {
            const int lo = g.nc[d] >= 3 ? -1 : 0, hi = g.nc[d] >= 2 ? 1 : 0;
            valid = valid && o[d] >= lo && o[d] <= hi;
            jj[d] = (cc[d] + o[d] + g.nc[d]) % g.nc[d];
            s_sh[t][d] = (double)((cc[d] + o[d]) - jj[d]) / g.nc[d] * g.L[d];
        }
        const int jc = (jj[2] * g.nc[1] + jj[1]) * g.nc[0] + jj[0];
        s_start[t] = start[jc];
        s_off[t + 1] = valid ? start[jc + 1] - start[jc] : 0;
    }
    __syncthreads();
    if (t == 0) {
        s_off[0] = 0;
        for (int k = 0; k < 27; k++) s_off[k + 1] += s_off[k];
    }
    __syncthreads();
    const int ncand = s_off[27];
    const int hs = start[c], he = start[c + 1];
    const float rc2_lo = (float)(rc2 * (1.0 - 2.0e-5)), rc2_hi = (float)(rc2 * (1.0 + 2.0e-5));
    for (int h0 = hs; h0 < he; h0 += kListThreads / G) {
        const int i = h0 + t / G, lane = t % G;
        const bool active = i < he;
        const double4 pi = active ? pos[i] : make_double4(0.0, 0.0, 0.0, 0.0);
        const float pfx = (float)pi.x, pfy = (float)pi.y, pfz = (float)pi.z;
        int count = 0;
        for (int r0 = 0; r0 < ncand; r0 += kListStage) {
            const int m = min(kListStage, ncand - r0);
            // stage candidates r0 .. r0+m-1 (skipped when a single round is
            // already in LDS from the previous pass of atoms)
            if (!(r0 == 0 && m == ncand && h0 > hs)) {
                __syncthreads();
                constexpr int U = 6;
                for (int k0 = t; k0 < m; k0 += U * kListThreads) {
                    double4 pk[U];
                    int jk[U], tk[U];
#pragma unroll
                    for (int u = 0; u < U; u++) {
                        const int k = k0 + u * kListThreads;
                        int q = 0;
                        if (k < m) {
                            const int kk = r0 + k;
                            while (s_off[q + 1] <= kk) q++;
                            jk[u] = s_start[q] + (kk - s_off[q]);
                            pk[u] = pos[jk[u]];
                        }
                        tk[u] = q;
                    }
#pragma unroll
                    for (int u = 0; u < U; u++) {
                        const int k = k0 + u * kListThreads;
                        if (k < m) {
                            const int q = tk[u];
                            cand[k] = make_float4((float)(pk[u].x + s_sh[q][0]), (float)(pk[u].y + s_sh[q][1]),
                                                  (float)(pk[u].z + s_sh[q][2]), __int_as_float(jk[u]));
                        }
                    }
                }
                __syncthreads();
            }
            if (!active) continue;
            if (small) {
                // 1-2 cells per axis: every candidate decided by exact minimum image
                for (int k = lane; k < m; k += G) {
                    const int j = __float_as_int(cand[k].w);
                    if (j == i) continue;
                    const double4 pj = pos[j];
                    const double dx = min_image_1(pi.x, pj.x, g.L[0], g.invL[0]);
                    const double dy = min_image_1(pi.y, pj.y, g.L[1], g.invL[1]);
                    const double dz = min_image_1(pi.z, pj.z, g.L[2], g.invL[2]);
                    const double rsq = __dadd_rn(__dadd_rn(__dmul_rn(dx, dx), __dmul_rn(dy, dy)), __dmul_rn(dz, dz));
                    if (rsq <= rc2) {
                        if (count < cap_lane) nbr[list_slot(i, count, lane, cap_lane)] = j;
                        count++;
                    }
                }
                continue;
            }
            // the accept test is predicated rather than branched (divergent
            // branches here cost more scalar than vector instructions), and
            // the rare band of pairs within 2e-5 of the cutoff is settled in
            // FP64 behind a wavefront-uniform test
#pragma unroll 2
            for (int k = lane; k < m; k += G) {
                const float4 e = cand[k];
                const int j = __float_as_int(e.w);
                const float ex = pfx - e.x, ey = pfy - e.y, ez = pfz - e.z;
                const float r2f = ex * ex + ey * ey + ez * ez;
                bool in = r2f < rc2_lo && j != i;
                const bool near = r2f >= rc2_lo && r2f <= rc2_hi && j != i;
                if (__any(near)) {
                    if (near) {
                        // x_i - (x_j + s L) with the stencil's s, which for a pair this
                        // close (well under L/2) is the minimum image
                        const double4 pj = pos[j];
                        const double dx = min_image_1(pi.x, pj.x, g.L[0], g.invL[0]);
                        const double dy = min_image_1(pi.y, pj.y, g.L[1], g.invL[1]);
                        const double dz = min_image_1(pi.z, pj.z, g.L[2], g.invL[2]);
                        const double rsq = __dadd_rn(__dadd_rn(__dmul_rn(dx, dx), __dmul_rn(dy, dy)),
                                                     __dmul_rn(dz, dz));
                        in = rsq <= rc2;
                    }
                }
                if (in) {
                    if (count < cap_lane) nbr[list_slot(i, count, lane, cap_lane)] = j;
                    count++;
                }
            }
        }
        if (active) {
            cnt[(size_t)i * G + lane] = count;
            if (count > cap_lane) atomicMax(&s_ovf, count);
        }
        // per-atom total (lanes of one atom are adjacent in the wavefront)
        int tot = count;
#pragma unroll
        for (int o = G / 2; o > 0; o >>= 1) tot += __shfl_xor(tot, o, 64);
        if (active && lane == 0) {
            atomicMax(&s_big, tot);
            atomicAdd(&s_tot, tot);
        }
    }
    // one global atomic per workgroup and flag: a single address hit by every
    // atom would serialise tens of thousands of atomics
    __syncthreads();
    if (t == 0) {
        if (s_ovf) atomicMax(&flags[0], s_ovf);
        atomicMax(&flags[1], s_big);
        atomicAdd(&flags[3], s_tot);
    }
}

// vnl_check (neighbours.F90:150-178): largest minimum-image displacement since
// the list was built, as the bit pattern of a non-negative double.
__global__ void k_max_disp(const double4 *pos, const double4 *xref, int n, CellGrid g,
                           unsigned long long *out)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    double r2 = 0.0;
    if (i < n) {
        const double4 p = pos[i], q = xref[i];
        const double dx = min_image_1(p.x, q.x, g.L[0], g.invL[0]);
        const double dy = min_image_1(p.y, q.y, g.L[1], g.invL[1]);
        const double dz = min_image_1(p.z, q.z, g.L[2], g.invL[2]);
        r2 = dx * dx + dy * dy + dz * dz;
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) r2 = fmax(r2, __shfl_xor(r2, o, 64));
    if ((threadIdx.x & 63) == 0) atomicMax(out, (unsigned long long)__double_as_longlong(r2));
}

static CellGrid make_grid(const Ctx *c)
{
    CellGrid g;
    for (int d = 0; d < 3; d++) {
        g.nc[d] = c->nc[d];
        g.L[d] = c->L[d];
        g.invL[d] = c->invL[d];
    }
    return g;
}

template <typename T>
static int grow(Ctx *c, T **p, size_t count)
{
    if (*p) (void)hipFree(*p);
    *p = nullptr;
    DLPX_HIP(c, hipMalloc((void **)p, sizeof(T) * (count > 0 ? count : 1)));
    return 0;
}

// Bin, sort, permute and build the list from c->pos (current order).
int build_list(Ctx *c)
{
    const int n = c->natms;
    const double rc_ext = c->prm.cutoff + c->prm.padding;
    const double cut = rc_ext + 1.0e-6;  // neighbours.F90:415 (smalldr)
    for (int d = 0; d < 3; d++) {
        if (rc_ext >= 0.5 * c->L[d])
            return fail(c, DLPX_ERR_CUTOFF, "cutoff+padding exceeds half the cell width");
        c->nc[d] = (int)(c->L[d] / cut);
        if (c->nc[d] < 1) return fail(c, DLPX_ERR_LINK_CELLS, "too few link cells");
    }
    const int ncell = c->nc[0] * c->nc[1] * c->nc[2];
    if (ncell + 1 > c->ncell_cap) {
        int r;
        if ((r = grow(c, &c->cell_count, ncell + 1))) return r;
        if ((r = grow(c, &c->cell_start, ncell + 1))) return r;
        if ((r = grow(c, &c->cell_fill, ncell + 1))) return r;
        c->ncell_cap = ncell + 1;
    }
    c->ncell = ncell;
    const CellGrid g = make_grid(c);
    hipStream_t s = c->stream;
    DLPX_HIP(c, hipMemsetAsync(c->cell_count, 0, sizeof(int32_t) * (ncell + 1), s));
    const int b = 256, nb = ceil_div(n, b);
    hipLaunchKernelGGL(k_wrap_bin, dim3(nb), dim3(b), 0, s, c->pos, n, g, c->cell_of, c->cell_count);
    hipLaunchKernelGGL(k_scan, dim3(1), dim3(1024), 0, s, c->cell_count, c->cell_start, c->cell_fill, ncell);
    hipLaunchKernelGGL(k_scatter, dim3(nb), dim3(b), 0, s, c->cell_of, c->cell_fill, c->perm, n);
    hipLaunchKernelGGL(k_sort_cells, dim3(ncell), dim3(256), 0, s, c->perm, c->cell_start, ncell);
    hipLaunchKernelGGL(k_permute, dim3(nb), dim3(b), 0, s, c->perm, n, c->pos, c->pos_alt, c->typ, c->typ_alt,
                       c->gid, c->gid_alt, c->resident ? c->vel : nullptr, c->resident ? c->vel_alt : nullptr,
                       c->slot, c->xref);
    hipLaunchKernelGGL(k_copy_back, dim3(nb), dim3(b), 0, s, n, c->pos_alt, c->pos, c->typ_alt, c->typ, c->gid_alt,
                       c->gid, c->resident ? c->vel_alt : nullptr, c->resident ? c->vel : nullptr);

    constexpr int G = kLanesPerAtom;
    for (int attempt = 0; attempt < 2; attempt++) {
        if (c->cap_lane == 0) {
            // first guess: density-based estimate of the per-atom count (neighbours.F90 max_list sizing)
            const double rho = (double)n / (c->L[0] * c->L[1] * c->L[2]);
            const double est = 4.18879020478639 * rho * rc_ext * rc_ext * rc_ext;  // 4/3 pi r^3 rho
            const int per_atom = c->prm.max_list > 0 ? c->prm.max_list : (int)(1.5 * est) + 64;
            c->cap_lane = (per_atom + G - 1) / G;
            int r;
            if ((r = grow(c, &c->nbr, (size_t)c->cap_lane * c->cap * G))) return r;
        }
        DLPX_HIP(c, hipMemsetAsync(c->flags, 0, sizeof(int32_t) * 4, s));
        if (c->timing) DLPX_HIP(c, hipEventRecord(c->ev[6], s));
        hipLaunchKernelGGL(k_build_list<G>, dim3(xcd_grid(ncell)), dim3(kListThreads), 0, s, c->pos, c->cap, g,
                           c->cell_start,
                           rc_ext * rc_ext, c->cap_lane, c->nbr, c->cnt, c->flags);
        if (c->timing) DLPX_HIP(c, hipEventRecord(c->ev[7], s));
        DLPX_HIP(c, hipMemcpyAsync(c->h_flags, c->flags, sizeof(int32_t) * 4, hipMemcpyDeviceToHost, s));
        DLPX_HIP(c, hipStreamSynchronize(s));
        if (c->timing) {
            float ms;
            DLPX_HIP(c, hipEventElapsedTime(&ms, c->ev[6], c->ev[7]));
            c->stage_ms[5] += ms;
        }
        if (c->h_flags[0] == 0) break;  // no lane overflowed
        if (c->prm.max_list > 0) {
            c->have_list = false;
            return fail(c, DLPX_ERR_LIST_OVERFLOW,
                        "neighbour list overflow: an atom has " + std::to_string(c->h_flags[1]) +
                            " neighbours, max_list is " + std::to_string(c->prm.max_list));
        }
        // grow with margin and rebuild
        c->cap_lane = c->h_flags[0] + c->h_flags[0] / 4 + 8;
        int r;
        if ((r = grow(c, &c->nbr, (size_t)c->cap_lane * c->cap * G))) return r;
    }
    if (c->h_flags[0] != 0) {
        c->have_list = false;
        return fail(c, DLPX_ERR_LIST_OVERFLOW, "neighbour list overflow after regrowth");
    }
    c->have_list = true;
    return 0;
}

int vnl_check_dev(Ctx *c, int *update)
{
    const int n = c->natms;
    hipStream_t s = c->stream;
    unsigned long long *d = (unsigned long long *)(c->flags + 4);
    DLPX_HIP(c, hipMemsetAsync(d, 0, sizeof(unsigned long long), s));
    hipLaunchKernelGGL(k_max_disp, dim3(ceil_div(n, 256)), dim3(256), 0, s, c->pos, c->xref, n, make_grid(c), d);
    DLPX_HIP(c, hipMemcpyAsync(c->h_flags + 4, d, sizeof(unsigned long long), hipMemcpyDeviceToHost, s));
    DLPX_HIP(c, hipStreamSynchronize(s));
    unsigned long long bits;
    memcpy(&bits, c->h_flags + 4, sizeof(bits));
    double r2;
    memcpy(&r2, &bits, sizeof(r2));
    const double tol = sqrt(r2);
    *update = (tol >= kHalfMinus * c->prm.padding) ? 1 : 0;
    return 0;
}

}  // namespace dlpx
