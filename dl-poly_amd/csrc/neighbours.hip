// Link-cell Verlet neighbour list on the GPU (neighbours.F90:356-1414).
//
// Atoms are binned into link cells of edge >= cutoff+padding or, with at
// least 3 link cells per axis, into the 2 x 2 x 2 subcells of each link cell
// (numbered cell-major, so a link cell's atoms stay contiguous), then
// counting-sorted by that key (deterministic order inside a key), and every
// per-atom array is permuted into that order, so the pair and SPME kernels
// read neighbours that sit close in HBM.  The list is a full list (each pair stored from both ends)
// split over G lanes per atom: lane g keeps every G-th candidate it accepts,
// element (i, c, g) lives at nbr[(i*cap_lane + c)*G + g] (list_slot), so one
// atom's entries, and one workgroup's atoms, are contiguous.
#include <cmath>
#include <cstring>
#include <vector>

#include "common.h"
#include "device_util.h"

namespace dlpx {

struct CellGrid {
    int nc[3];
    double L[3], invL[3];
    int sub;  // 1: bin by link cell; 2: by subcell (half the link-cell edge)
};

// Binning key of a subcell (sx, sy, sz) of edge L/(2 nc): link-cell-major,
// then the 2 x 2 x 2 subcells of that link cell, so a link cell's atoms are
// one contiguous run of the sorted order.
__host__ __device__ __forceinline__ int subcell_key(const CellGrid &g, int sx, int sy, int sz)
{
    const int blk = ((sz >> 1) * g.nc[1] + (sy >> 1)) * g.nc[0] + (sx >> 1);
    return blk * 8 + ((sz & 1) * 4 + (sy & 1) * 2 + (sx & 1));
}

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

// Every list-build kernel takes `upd`: null on the host-decided path, else
// the device word that stage 1 set (F_UPD), and a zero there means no
// rebuild is due, so the kernel returns at once.  That lets a captured step
// graph hold the whole rebuild without a host round trip per step.
__device__ __forceinline__ bool skip_build(const int32_t *upd) { return upd && *upd == 0; }

// The sort before a list build runs as three kernels:
//   k_wrap_scan     fold positions into the primary cell (relocate_particles /
//                   pbcshift), bin them, and in the last workgroup to finish
//                   turn the counts into start offsets;
//   k_scatter_copy  place each atom's index in its key's run, and copy the
//                   atom data to the spare arrays;
//   k_sort_permute  order each key's run by the previous index and gather the
//                   atoms from the spare arrays into the primary ones.
// The primary arrays keep their addresses (captured step graphs hold them).

// Exclusive scan of the per-key counts by one workgroup of 1024 threads, each
// thread taking kScanRun consecutive keys per pass (subcell binning has
// hundreds of thousands of keys at a million atoms).  The counts are zeroed
// once read, ready for the next build.  With subcells (cell_start not null)
// the link-cell starts are the starts of every 8th key.
constexpr int kScanRun = 8;
constexpr int kWrapThreads = 1024;
__device__ __forceinline__ void scan_keys(int32_t *count, int32_t *start, int32_t *fill, int n, int32_t *cell_start)
{
    __shared__ int32_t part[kWrapThreads];
    __shared__ int32_t carry;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    for (int base = 0; base < n; base += kWrapThreads * kScanRun) {
        const int i0 = base + threadIdx.x * kScanRun;
        int v[kScanRun], sum = 0;
#pragma unroll
        for (int k = 0; k < kScanRun; k++) {
            v[k] = i0 + k < n ? load_coherent(&count[i0 + k]) : 0;
            sum += v[k];
        }
#pragma unroll
        for (int k = 0; k < kScanRun; k++)
            if (i0 + k < n) count[i0 + k] = 0;
        part[threadIdx.x] = sum;
        __syncthreads();
        for (int o = 1; o < kWrapThreads; o <<= 1) {
            int t = threadIdx.x >= o ? part[threadIdx.x - o] : 0;
            __syncthreads();
            part[threadIdx.x] += t;
            __syncthreads();
        }
        int run = carry + part[threadIdx.x] - sum;
#pragma unroll
        for (int k = 0; k < kScanRun; k++) {
            if (i0 + k < n) {
                start[i0 + k] = run;
                fill[i0 + k] = run;
                if (cell_start && ((i0 + k) & 7) == 0) cell_start[(i0 + k) >> 3] = run;
            }
            run += v[k];
        }
        __syncthreads();
        if (threadIdx.x == kWrapThreads - 1) carry += part[kWrapThreads - 1];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        start[n] = carry;
        if (cell_start) cell_start[n >> 3] = carry;
    }
}

struct WrapArgs {
    double4 *pos;
    int n;
    CellGrid g;
    int32_t *cell_of, *count;
    int32_t *start, *fill;   // key starts (and the scatter's fill pointers)
    int nkey;
    int32_t *cell_start;     // subcells: link-cell starts; else null (start is them)
    unsigned int *done;      // workgroups finished (zero between builds)
    const int32_t *upd;
};
__device__ __forceinline__ void wrap_scan_body(const WrapArgs &a, int bid, int nb)
{
    __shared__ bool last;
    const CellGrid &g = a.g;
    const int i = bid * blockDim.x + threadIdx.x;
    if (i < a.n) {
        double4 p = a.pos[i];
        p.x -= g.L[0] * floor(p.x * g.invL[0] + 0.5);
        p.y -= g.L[1] * floor(p.y * g.invL[1] + 0.5);
        p.z -= g.L[2] * floor(p.z * g.invL[2] + 0.5);
        // floor(x/L + 0.5) can leave x == L/2 by rounding
        if (p.x >= 0.5 * g.L[0]) p.x -= g.L[0];
        if (p.y >= 0.5 * g.L[1]) p.y -= g.L[1];
        if (p.z >= 0.5 * g.L[2]) p.z -= g.L[2];
        a.pos[i] = p;
        const int cx = cell_coord(p.x, g.L[0], g.invL[0], g.sub * g.nc[0]);
        const int cy = cell_coord(p.y, g.L[1], g.invL[1], g.sub * g.nc[1]);
        const int cz = cell_coord(p.z, g.L[2], g.invL[2], g.sub * g.nc[2]);
        const int c = g.sub == 2 ? subcell_key(g, cx, cy, cz) : (cz * g.nc[1] + cy) * g.nc[0] + cx;
        a.cell_of[i] = c;
        // atoms are mostly in cell order already: one atomic per run of equal
        // cells within the wavefront
        int head, len;
        run_of(c, head, len);
        if ((int)(threadIdx.x & 63) == head) atomicAdd(&a.count[c], len);
    }
    // the last workgroup to finish scans the counts (device-scope atomics,
    // read coherently)
    wait_memory();
    __syncthreads();
    if (threadIdx.x == 0) last = atomicAdd(a.done, 1u) == (unsigned)nb - 1;
    __syncthreads();
    if (!last) return;
    scan_keys(a.count, a.start, a.fill, a.nkey, a.cell_start);
    if (threadIdx.x == 0) __hip_atomic_store(a.done, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

struct ScatterArgs {
    const int32_t *cell_of;
    int32_t *fill, *perm;
    int n;
    const int32_t *upd;
    int32_t *flags;  // F_ERR on an inconsistent count
    // the atom data, copied to the spare arrays for k_sort_permute to gather from
    const double4 *pos;
    double4 *pos_alt;
    const int32_t *typ;
    int32_t *typ_alt;
    const int32_t *gid;
    int32_t *gid_alt;
    const double4 *vel;  // may be null
    double4 *vel_alt;
};
__device__ __forceinline__ void scatter_copy_body(const ScatterArgs &a, int bid)
{
    const int i = bid * blockDim.x + threadIdx.x;
    if (i >= a.n) return;
    a.pos_alt[i] = a.pos[i];
    a.typ_alt[i] = a.typ[i];
    a.gid_alt[i] = a.gid[i];
    if (a.vel) a.vel_alt[i] = a.vel[i];
    const int c = a.cell_of[i];
    int head, len;
    run_of(c, head, len);
    const int lane = threadIdx.x & 63;
    int base = 0;
    if (lane == head) base = atomicAdd(&a.fill[c], len);
    base = __shfl(base, head, 64);
    const int dst = base + (lane - head);
    if ((unsigned)dst < (unsigned)a.n) a.perm[dst] = i;
    else a.flags[F_ERR] = 1;
}

// Make the order inside each key deterministic (ascending previous index):
// one workgroup per key, every member's rank is the number of smaller
// members (indices are distinct), so the sort is one parallel pass in LDS;
// then the workgroup gathers its atoms from the spare arrays.  A key with
// more than kSortLds members (a very crowded cell) is sorted by one thread in
// global memory.
constexpr int kSortLds = 2048;
struct SortArgs {
    int32_t *perm;
    const int32_t *start;
    int nkey, n;
    const int32_t *upd;
    int32_t *flags;  // F_ERR on an index outside the atoms
    const double4 *pos_alt;
    double4 *pos;
    const int32_t *typ_alt;
    int32_t *typ;
    const int32_t *gid_alt;
    int32_t *gid;
    const double4 *vel_alt;  // may be null
    double4 *vel;
    int32_t *slot;
    double4 *xref;
};
__device__ __forceinline__ void permute_one(const SortArgs &a, int i, int o)
{
    // perm and gid hold indices below n by construction; the test keeps a
    // broken count from turning into stray memory accesses, and is reported
    // as an internal error
    if ((unsigned)o >= (unsigned)a.n) {
        a.flags[F_ERR] = 1;
        return;
    }
    const double4 p = a.pos_alt[o];
    a.pos[i] = p;
    a.xref[i] = p;
    a.typ[i] = a.typ_alt[o];
    const int g = a.gid_alt[o];
    a.gid[i] = g;
    if ((unsigned)g < (unsigned)a.n) a.slot[g] = i;
    else a.flags[F_ERR] = 1;
    if (a.vel) a.vel[i] = a.vel_alt[o];
}
__device__ __forceinline__ void sort_permute_body(const SortArgs &a, int c)
{
    __shared__ int32_t v[kSortLds], sorted[kSortLds];
    int32_t *perm = a.perm;
    const int s = a.start[c], n = a.start[c + 1] - s;
    if (n <= 0) return;
    if (n > kSortLds) {
        if (threadIdx.x == 0) {
            for (int q = s + 1; q < s + n; q++) {
                const int x = perm[q];
                int b = q - 1;
                while (b >= s && perm[b] > x) {
                    perm[b + 1] = perm[b];
                    b--;
                }
                perm[b + 1] = x;
            }
        }
        // the sorted run was written by thread 0: make it visible to the
        // workgroup past its L1
        __threadfence();
        __syncthreads();
        for (int k = threadIdx.x; k < n; k += blockDim.x)
            permute_one(a, s + k, __hip_atomic_load(&perm[s + k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
        return;
    }
    for (int k = threadIdx.x; k < n; k += blockDim.x) v[k] = perm[s + k];
    __syncthreads();
    for (int k = threadIdx.x; k < n; k += blockDim.x) {
        const int x = v[k];
        int rank = 0;
        for (int j = 0; j < n; j++) rank += v[j] < x;
        sorted[rank] = x;
        perm[s + rank] = x;
    }
    __syncthreads();
    for (int k = threadIdx.x; k < n; k += blockDim.x) permute_one(a, s + k, sorted[k]);
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
// The workgroup's partial counters into the build's flags (one global
// atomic each: a single address hit by every atom would serialise tens of
// thousands of atomics); an overflow also sets the sticky F_OVF_ANY that a
// resident run checks once at its end.
__device__ __forceinline__ void build_flags(int32_t *flags, int ovf, int big, int tot)
{
    if (ovf) {
        atomicMax(&flags[F_OVF], ovf);
        atomicMax(&flags[F_OVF_ANY], ovf);
    }
    atomicMax(&flags[F_IBIG], big);
    atomicAdd(&flags[F_PAIRS], tot);
}

struct BuildArgs {
    const double4 *pos;
    CellGrid g;
    const int32_t *start;  // link-cell starts (k_build_list) or subcell-key starts (k_build_list_sub)
    double rc2;
    int cap_lane;
    int32_t *nbr, *cnt, *flags;
    const int32_t *upd;
    int32_t *sub_of;  // subcell build: each atom's subcell (see Ctx::sub_of)
};
template <int G>
__device__ __forceinline__ void build_list_body(const BuildArgs &a, int c)
{
    static_assert(G == kLanesPerAtom, "list_slot assumes kLanesPerAtom lanes per atom");
    __shared__ float4 cand[kListStage];
    __shared__ int s_start[27], s_off[28];
    __shared__ double s_sh[27][3];
    __shared__ int s_ovf, s_big, s_tot;  // workgroup partials of flags[0], [1], [3]
    const double4 *__restrict__ pos = a.pos;
    const CellGrid &g = a.g;
    const int32_t *__restrict__ start = a.start;
    const double rc2 = a.rc2;
    const int cap_lane = a.cap_lane;
    int32_t *__restrict__ nbr = a.nbr;
    int32_t *__restrict__ cnt = a.cnt;
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
    __syncthreads();
    if (t == 0) build_flags(a.flags, s_ovf, s_big, s_tot);
}

// List build over subcells of half the link-cell edge (used when every axis
// has at least 3 link cells).  One workgroup per link cell (2 x 2 x 2
// subcells); the 6 x 6 x 6 subcells around it are staged in LDS as float4
// {x + sx, y + sy, z + sz, j}, grouped by subcell in x-fastest order, so that
// for an atom the 5 subcells of each of its 25 (y, z) rows are one
// contiguous run of candidates.  An atom is never more than 2 subcells from
// a neighbour (the subcell edge is at least half of cutoff+padding), so each
// atom tests 125 subcells' atoms instead of the 27 link cells' (about 4.6
// link cells' worth of candidates instead of 27).
//
// One atom per wavefront at a time: the 64 lanes test 64 candidates of a run,
// and the accepted ones are packed with a ballot into the atom's next list
// entries, so the loops and the list counter are wavefront-uniform.  Entry e
// of atom i sits at list_slot(i, e / G, e % G), the layout of the link-cell
// build, and list lane g then holds entries e = g, g + G, ...  The accept
// test (FP32 filter, FP64 decision within 2e-5 of the cutoff) is the link-cell
// build's, so both give the same set.
constexpr int kSubSide = 6, kSubStencil = kSubSide * kSubSide * kSubSide;
constexpr int kSubAtomChunk = 256;  // atoms of a link cell whose counts persist across staging rounds
template <int G>
__device__ __forceinline__ void build_list_sub_body(const BuildArgs &a, int b)
{
    static_assert(G == kLanesPerAtom, "list_slot assumes kLanesPerAtom lanes per atom");
    __shared__ float4 cand[kListStage];
    __shared__ int s_start[kSubStencil], s_off[kSubStencil + 1];  // s_off: sizes, then their exclusive scan
    __shared__ double s_sh[3][kSubSide];
    __shared__ int s_lo[9];
    __shared__ int s_acnt[kSubAtomChunk];
    __shared__ int s_ovf, s_big, s_tot;
    const double4 *__restrict__ pos = a.pos;
    const CellGrid &g = a.g;
    const int32_t *__restrict__ key_start = a.start;
    const double rc2 = a.rc2;
    const int cap_lane = a.cap_lane;
    int32_t *__restrict__ nbr = a.nbr;
    int32_t *__restrict__ cnt = a.cnt;
    int32_t *__restrict__ sub_of = a.sub_of;
    if (b >= g.nc[0] * g.nc[1] * g.nc[2]) return;
    const int t = threadIdx.x, lane = t & 63, wid = t >> 6;
    constexpr int NW = kListThreads / 64;
    const int bc[3] = {b % g.nc[0], (b / g.nc[0]) % g.nc[1], b / (g.nc[0] * g.nc[1])};
    if (t == 0) {
        s_ovf = 0;
        s_big = 0;
        s_tot = 0;
    }
    if (t < kSubStencil) {
        int sc[3];
        const int q[3] = {t % kSubSide, (t / kSubSide) % kSubSide, t / (kSubSide * kSubSide)};
#pragma unroll
        for (int d = 0; d < 3; d++) {
            const int ns = 2 * g.nc[d], u = 2 * bc[d] - 2 + q[d];
            sc[d] = (u + ns) % ns;  // ns >= 6: the 6 stencil subcells of an axis are distinct
        }
        const int key = subcell_key(g, sc[0], sc[1], sc[2]);
        s_start[t] = key_start[key];
        s_off[t] = key_start[key + 1] - key_start[key];
    } else if (t < kSubStencil + 3 * kSubSide) {
        const int d = (t - kSubStencil) / kSubSide, qd = (t - kSubStencil) % kSubSide;
        const int ns = 2 * g.nc[d], u = 2 * bc[d] - 2 + qd, su = (u + ns) % ns;
        s_sh[d][qd] = (double)(u - su) / ns * g.L[d];  // image shift: -L, 0 or +L
    } else if (t < kSubStencil + 3 * kSubSide + 9) {
        s_lo[t - kSubStencil - 3 * kSubSide] = key_start[8 * b + (t - kSubStencil - 3 * kSubSide)];
    }
    __syncthreads();
    if (t < 64) {
        // exclusive scan of the 216 subcell sizes in place: 4 per lane (each
        // lane reads its own entries before writing them), then across lanes
        int v[4], sum = 0;
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const int q = 4 * t + k;
            v[k] = q < kSubStencil ? s_off[q] : 0;
            sum += v[k];
        }
        int incl = sum;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const int y = __shfl_up(incl, o, 64);
            if (t >= o) incl += y;
        }
        int run = incl - sum;
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const int q = 4 * t + k;
            if (q < kSubStencil) s_off[q] = run;
            run += v[k];
        }
        if (t == 63) s_off[kSubStencil] = incl;
    }
    __syncthreads();
    const int ncand = s_off[kSubStencil];
    const int hs = s_lo[0], nat = s_lo[8] - hs;
    const int cap_tot = cap_lane * G;
    const float rc2_lo = (float)(rc2 * (1.0 - 2.0e-5)), rc2_hi = (float)(rc2 * (1.0 + 2.0e-5));
    const unsigned long long lt_mask = (1ull << lane) - 1ull;
    for (int a0 = 0; a0 < nat; a0 += kSubAtomChunk) {
        const int a1 = min(nat, a0 + kSubAtomChunk);
        for (int r0 = 0; r0 < ncand; r0 += kListStage) {
            const int m = min(kListStage, ncand - r0);
            const bool last = r0 + m == ncand;
            // stage candidates r0 .. r0+m-1 (kept when one round holds them all)
            if (!(r0 == 0 && last && a0 > 0)) {
                __syncthreads();
                constexpr int U = 6;
                for (int k0 = t; k0 < m; k0 += U * kListThreads) {
                    double4 pk[U];
                    int jk[U], qk[U];
#pragma unroll
                    for (int u = 0; u < U; u++) {
                        const int k = k0 + u * kListThreads;
                        qk[u] = 0;
                        if (k < m) {
                            const int kk = r0 + k;
                            int lo = 0, hi = kSubStencil;  // s_off[lo] <= kk < s_off[hi]
                            while (hi - lo > 1) {
                                const int mid = (lo + hi) >> 1;
                                if (s_off[mid] <= kk) lo = mid;
                                else hi = mid;
                            }
                            qk[u] = lo;
                            jk[u] = s_start[lo] + (kk - s_off[lo]);
                            pk[u] = pos[jk[u]];
                        }
                    }
#pragma unroll
                    for (int u = 0; u < U; u++) {
                        const int k = k0 + u * kListThreads;
                        if (k < m) {
                            const int q = qk[u];
                            const int qx = q % kSubSide, qy = (q / kSubSide) % kSubSide, qz = q / (kSubSide * kSubSide);
                            // w: j and the candidate's staged x subcell (for the entry's code)
                            cand[k] = make_float4((float)(pk[u].x + s_sh[0][qx]), (float)(pk[u].y + s_sh[1][qy]),
                                                  (float)(pk[u].z + s_sh[2][qz]),
                                                  __int_as_float(jk[u] | (qx << kListCodeShift)));
                        }
                    }
                }
                __syncthreads();
            }
            // the next atom's position is loaded while this atom is processed
            int a = a0 + wid;
            double4 pnext = a < a1 ? pos[hs + a] : make_double4(0.0, 0.0, 0.0, 0.0);
            for (; a < a1; a += NW) {
                const int i = hs + a;
                const double4 pi = pnext;
                if (a + NW < a1) pnext = pos[i + NW];
                // the atom's subcell among the link cell's 8 (s_lo is non-decreasing)
                int l = 0;
#pragma unroll
                for (int k = 1; k < 8; k++) l += s_lo[k] <= i;
                const int lx = l & 1, ly = (l >> 1) & 1, lz = l >> 2;
                if (r0 == 0 && lane == 0)
                    sub_of[i] = (2 * bc[0] + lx) | ((2 * bc[1] + ly) << 10) | ((2 * bc[2] + lz) << 20);
                const float pfx = (float)pi.x, pfy = (float)pi.y, pfz = (float)pi.z;
                // lane r < 25 holds the staged candidate range of row r (its 5
                // subcells in x at dz = r / 5, dy = r % 5), relative to the round
                int rlo = 0, rhi = 0;
                if (lane < 25) {
                    const int qrow = ((lz + lane / 5) * kSubSide + (ly + lane % 5)) * kSubSide + lx;
                    rlo = max(s_off[qrow], r0) - r0;
                    rhi = min(s_off[qrow + 5], r0 + m) - r0;
                }
                int count = __builtin_amdgcn_readfirstlane(r0 == 0 ? 0 : s_acnt[a - a0]);
                int32_t *const row_out = nbr + list_slot(i, 0, 0, cap_lane);
                // one chunk of up to 64 candidates; accepted ones are packed into the
                // next list entries (an overflowing list is rebuilt, so entries past
                // the capacity all land on its last slot)
                // an accepted candidate of row `row` is stored as j | code << 25, its
                // subcell offset from i's: dz + 2 = row / 5, dy + 2 = row % 5,
                // dx + 2 = (staged x subcell) - lx
                auto test = [&](const float4 e, const bool v, const int row) {
                    const int w = __float_as_int(e.w);
                    const int j = list_j(w);
                    const float ex = pfx - e.x, ey = pfy - e.y, ez = pfz - e.z;
                    const float r2f = ex * ex + ey * ey + ez * ez;
                    bool in = v && r2f < rc2_lo && j != i;
                    const bool near = v && r2f >= rc2_lo && r2f <= rc2_hi && j != i;
                    if (__any(near)) {
                        if (near) {
                            // x_i - (x_j + s L) with the stencil's s is the minimum
                            // image for a pair this close (well under L/2)
                            const double4 pj = pos[j];
                            const double dx = min_image_1(pi.x, pj.x, g.L[0], g.invL[0]);
                            const double dy = min_image_1(pi.y, pj.y, g.L[1], g.invL[1]);
                            const double dz = min_image_1(pi.z, pj.z, g.L[2], g.invL[2]);
                            const double rsq = __dadd_rn(__dadd_rn(__dmul_rn(dx, dx), __dmul_rn(dy, dy)),
                                                         __dmul_rn(dz, dz));
                            in = rsq <= rc2;
                        }
                    }
                    const unsigned long long acc = __ballot(in);
                    const int pre = __builtin_amdgcn_mbcnt_hi((unsigned)(acc >> 32),
                                                              __builtin_amdgcn_mbcnt_lo((unsigned)acc, 0u));
                    // w = j | qx << 25, so the entry j | (row 5 + qx - lx) << 25 is w
                    // plus a row constant (modulo 2^32; the code stays below 128)
                    if (in)
                        row_out[min(count + pre, cap_tot - 1)] =
                            (int32_t)((uint32_t)w + ((uint32_t)(row * 5 - lx) << kListCodeShift));
                    count += __popcll(acc);
                };
                // five rows at a time: their first chunks' LDS reads are issued together
#pragma unroll 1
                for (int g5 = 0; g5 < 25; g5 += 5) {
                    int lo[5], hi[5];
                    bool v[5];
                    float4 e[5];
#pragma unroll
                    for (int u = 0; u < 5; u++) {
                        lo[u] = __builtin_amdgcn_readlane(rlo, g5 + u);
                        hi[u] = __builtin_amdgcn_readlane(rhi, g5 + u);
                        v[u] = lo[u] + lane < hi[u];
                        e[u] = cand[v[u] ? lo[u] + lane : 0];
                    }
#pragma unroll
                    for (int u = 0; u < 5; u++) test(e[u], v[u], g5 + u);
#pragma unroll
                    for (int u = 0; u < 5; u++)
                        for (int k0 = lo[u] + 64; k0 < hi[u]; k0 += 64) {
                            const bool vv = k0 + lane < hi[u];
                            test(cand[vv ? k0 + lane : 0], vv, g5 + u);
                        }
                }
                if (!last) {
                    if (lane == 0) s_acnt[a - a0] = count;
                    continue;
                }
                // per-lane counts of the G-way interleave: lane g holds entries g, g+G, ...
                if (lane < G) cnt[(size_t)i * G + lane] = (count - lane + G - 1) / G;
                if (lane == 0) {
                    if (count > cap_tot) atomicMax(&s_ovf, (count + G - 1) / G);
                    atomicMax(&s_big, count);
                    atomicAdd(&s_tot, count);
                }
            }
        }
    }
    __syncthreads();
    if (t == 0) build_flags(a.flags, s_ovf, s_big, s_tot);
}

// Listed pairs closer than r (each pair is met from both ends: the count is
// halved by the caller), FP64 minimum image; thread t = i*G + g walks lane g
// of sorted atom i.
__global__ void __launch_bounds__(256) k_count_pairs(const double4 *__restrict__ pos, const int32_t *__restrict__ nbr,
                                                     const int32_t *__restrict__ cnt, int n, int cap_lane, double r2,
                                                     CellGrid g, unsigned long long *out)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int i = t / kLanesPerAtom, lane = t % kLanesPerAtom;
    unsigned long long k = 0;
    if (i < n) {
        const double4 pi = pos[i];
        const int m = min(cnt[t], cap_lane);
        for (int e = 0; e < m; e++) {
            const double4 pj = pos[list_j(nbr[list_slot(i, e, lane, cap_lane)])];
            const double dx = min_image_1(pi.x, pj.x, g.L[0], g.invL[0]);
            const double dy = min_image_1(pi.y, pj.y, g.L[1], g.invL[1]);
            const double dz = min_image_1(pi.z, pj.z, g.L[2], g.invL[2]);
            k += (dx * dx + dy * dy + dz * dz < r2) ? 1ull : 0ull;
        }
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) k += __shfl_xor(k, o, 64);
    if ((threadIdx.x & 63) == 0 && k) atomicAdd(out, k);
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

// Frozen-frozen and excluded pairs leave the list, as link_cell_pairs moves
// them past list(0,i) so that the two-body loop never sees them
// (neighbours.F90:1198-1250: a pair of two frozen atoms, lfrzn /= 0, and
// every j found in list_excl(:,i)).  Thread t = i*G + g compacts lane g of
// sorted atom i in place, keeping the order of what stays; F_PAIRS (the
// entry count) loses what was dropped.  excl_off / frz may be null.
struct ExclListArgs {
    int n, cap_lane;
    int32_t *nbr, *cnt;
    const int32_t *gid, *excl_off, *excl_ids, *frz;
    int32_t *flags;
    const int32_t *upd;
};
__device__ __forceinline__ void list_exclude_body(const ExclListArgs &a, int bid)
{
    constexpr int G = kLanesPerAtom;
    const int t = bid * blockDim.x + threadIdx.x;
    const int i = t / G, g = t % G;
    int dropped = 0;
    if (i < a.n) {
        const int h = a.gid[i];
        const int e0 = a.excl_off ? a.excl_off[h] : 0, e1 = a.excl_off ? a.excl_off[h + 1] : 0;
        const bool fi = a.frz && a.frz[h] != 0;
        if (e1 > e0 || fi) {
            const int m = min(a.cnt[t], a.cap_lane);
            int w = 0;
            for (int k = 0; k < m; k++) {
                const int32_t en = a.nbr[list_slot(i, k, g, a.cap_lane)];
                const int hj = a.gid[list_j(en)];
                bool ex = fi && a.frz[hj] != 0;
                for (int e = e0; e < e1; e++) ex = ex || a.excl_ids[e] == hj;
                if (!ex) a.nbr[list_slot(i, w++, g, a.cap_lane)] = en;
            }
            dropped = m - w;
            if (a.cnt[t] <= a.cap_lane) a.cnt[t] = w;  // an overflowed lane is rebuilt anyway
        }
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) dropped += __shfl_xor(dropped, o, 64);
    if ((threadIdx.x & 63) == 0 && dropped) atomicSub(&a.flags[F_PAIRS], dropped);
}

// Single-system kernels (arguments by value) and batched ones (argument
// structs from device memory, one per system, see batch_find) around the
// bodies above; every one first honours the rebuild gate.  A batched launch
// may also take one gate for the whole batch as a kernel argument (a batch
// of one system: its F_UPD), which a skipped launch reads without first
// fetching its argument struct.  A batch of several systems runs the
// compacted form (_cbatch): a fixed grid of workgroups loops over the blocks
// of only the systems whose rebuild is due, from the plan k_list_plan wrote
// (their indices act[], count at [nsys], and per kernel `row` the prefix sums
// of their block counts), so the systems without a rebuild cost no dispatch.
#define DLPX_LIST_KERNELS(NAME, ARGS, NT, CALL_SINGLE, CALL_BATCH, XCD)                                      \
    __global__ void __launch_bounds__(NT) NAME(ARGS a)                                                    \
    {                                                                                                     \
        if (skip_build(a.upd)) return;                                                                    \
        CALL_SINGLE;                                                                                      \
    }                                                                                                     \
    __global__ void __launch_bounds__(NT) NAME##_batch(const ARGS *__restrict__ av,                       \
                                                       const int32_t *__restrict__ boff, int nsys,        \
                                                       const int32_t *__restrict__ gate)                  \
    {                                                                                                     \
        if (skip_build(gate)) return;                                                                     \
        BatchPos bp;                                                                                      \
        if (XCD) {                                                                                        \
            if (!batch_pos(boff, nsys, bp)) return;                                                       \
        } else {                                                                                          \
            if ((int)blockIdx.x >= boff[nsys]) return;                                                    \
            bp.s = batch_find(boff, nsys, blockIdx.x);                                                    \
            bp.b = blockIdx.x - boff[bp.s];                                                               \
            bp.nb = boff[bp.s + 1] - boff[bp.s];                                                          \
        }                                                                                                 \
        const ARGS &a = av[bp.s];                                                                          \
        if (skip_build(a.upd)) return;                                                                    \
        CALL_BATCH;                                                                                       \
    }                                                                                                     \
    __global__ void __launch_bounds__(NT) NAME##_cbatch(const ARGS *__restrict__ av,                      \
                                                        const int32_t *__restrict__ plan, int nsys, int row) \
    {                                                                                                     \
        const int32_t *act = plan, *coff = plan + nsys + 1 + row * (nsys + 1);                            \
        const int nact = plan[nsys], tot = coff[nact];                                                    \
        for (int lb = blockIdx.x; lb < tot; lb += gridDim.x) {                                            \
            int lo = 0, hi = nact; /* coff[lo] <= lb < coff[hi] */                                       \
            while (hi - lo > 1) {                                                                         \
                const int mid = (lo + hi) >> 1;                                                           \
                if (coff[mid] <= lb) lo = mid;                                                            \
                else hi = mid;                                                                            \
            }                                                                                             \
            BatchPos bp;                                                                                  \
            bp.s = act[lo];                                                                               \
            bp.b = lb - coff[lo];                                                                         \
            bp.nb = coff[lo + 1] - coff[lo];                                                              \
            const ARGS &a = av[bp.s];                                                                      \
            CALL_BATCH;                                                                                   \
            __syncthreads(); /* the next block's use of LDS */                                           \
        }                                                                                                 \
    }
DLPX_LIST_KERNELS(k_wrap_scan, WrapArgs, kWrapThreads, wrap_scan_body(a, blockIdx.x, gridDim.x),
                  wrap_scan_body(a, bp.b, bp.nb), false)
DLPX_LIST_KERNELS(k_scatter_copy, ScatterArgs, 256, scatter_copy_body(a, blockIdx.x), scatter_copy_body(a, bp.b), false)
DLPX_LIST_KERNELS(k_sort_permute, SortArgs, 256, sort_permute_body(a, blockIdx.x), sort_permute_body(a, bp.b), false)
DLPX_LIST_KERNELS(k_list_exclude, ExclListArgs, 256, list_exclude_body(a, blockIdx.x), list_exclude_body(a, bp.b),
                  false)
DLPX_LIST_KERNELS(k_build_list, BuildArgs, kListThreads,
                  build_list_body<kLanesPerAtom>(a, xcd_block(blockIdx.x, gridDim.x)),
                  build_list_body<kLanesPerAtom>(a, bp.b), true)
DLPX_LIST_KERNELS(k_build_list_sub, BuildArgs, kListThreads,
                  build_list_sub_body<kLanesPerAtom>(a, xcd_block(blockIdx.x, gridDim.x)),
                  build_list_sub_body<kLanesPerAtom>(a, bp.b), true)
#undef DLPX_LIST_KERNELS

// The plan of a compacted list chain: the systems whose F_UPD asks for a
// rebuild (act[0..nact), nact at [nsys]) and, per kernel row q, the prefix
// sums coff_q[k] of their block counts nb[q][act[k]].  One workgroup of 256
// threads, one system per thread per round: the compaction and the six
// prefix sums are workgroup scans (a serial loop over the systems took 46 us
// for 256 of them).
constexpr int kListKernels = 6;  // wrap, scatter, sort, subcell build, link-cell build, exclusions
__device__ __forceinline__ int block_excl_scan256(int v, int *wsum, int &total)
{
    // exclusive scan of one value per thread over 256 threads
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    int incl = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int y = __shfl_up(incl, o, 64);
        if (lane >= o) incl += y;
    }
    if (lane == 63) wsum[w] = incl;
    __syncthreads();
    int off = 0;
    for (int k = 0; k < w; k++) off += wsum[k];
    total = wsum[0] + wsum[1] + wsum[2] + wsum[3];
    __syncthreads();
    return off + incl - v;
}
__global__ void __launch_bounds__(256) k_list_plan(int nsys, int32_t *const *__restrict__ gates,
                                                   const int32_t *__restrict__ nb, int32_t *__restrict__ plan)
{
    __shared__ int wsum[4];
    const int t = threadIdx.x;
    int32_t *act = plan, *coff = plan + nsys + 1;
    int nact = 0, run[kListKernels] = {0, 0, 0, 0, 0, 0};
    for (int base = 0; base < nsys; base += 256) {
        const int s = base + t;
        const int f = s < nsys && *gates[s] != 0 ? 1 : 0;
        int tot;
        const int k = nact + block_excl_scan256(f, wsum, tot);
        if (f) act[k] = s;
#pragma unroll
        for (int q = 0; q < kListKernels; q++) {
            const int v = f ? nb[q * nsys + s] : 0;
            int tq;
            const int e = block_excl_scan256(v, wsum, tq);
            if (f) coff[q * (nsys + 1) + k] = run[q] + e;
            run[q] += tq;
        }
        nact += tot;
    }
    if (t == 0) {
        for (int q = 0; q < kListKernels; q++) coff[q * (nsys + 1) + nact] = run[q];
        act[nsys] = nact;
    }
}

static CellGrid make_grid(const Ctx *c)
{
    CellGrid g;
    for (int d = 0; d < 3; d++) {
        g.nc[d] = c->nc[d];
        g.L[d] = c->L[d];
        g.invL[d] = c->invL[d];
    }
    g.sub = 1;
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

// The arguments of every kernel of one list build of a context, and its
// launch shapes.  Sizes the cell arrays (so it is never called inside a
// stream capture before an eager build has sized them).
struct ListChain {
    WrapArgs wrap;
    ScatterArgs scat;
    SortArgs sort;
    BuildArgs build;
    ExclListArgs excl;
    bool sub, excl_on;
    int n, ncell, nkey;
};

static int list_chain(Ctx *c, const int32_t *upd, ListChain &lc)
{
    const int n = c->natms;
    if (c->excl_total > 0 && c->excl_natms != n)
        return fail(c, DLPX_ERR_STATE, "the exclusion lists were given for " + std::to_string(c->excl_natms) +
                                           " atoms, the configuration has " + std::to_string(n));
    if (c->nfrozen > 0 && c->frz_natms != n)
        return fail(c, DLPX_ERR_STATE, "the frozen flags were given for " + std::to_string(c->frz_natms) +
                                           " atoms, the configuration has " + std::to_string(n));
    const double rc_ext = c->prm.cutoff + c->prm.padding;
    const double cut = rc_ext + 1.0e-6;  // neighbours.F90:415 (smalldr)
    for (int d = 0; d < 3; d++) {
        if (rc_ext >= 0.5 * c->L[d])
            return fail(c, DLPX_ERR_CUTOFF, "cutoff+padding exceeds half the cell width");
        c->nc[d] = (int)(c->L[d] / cut);
        if (c->nc[d] < 1) return fail(c, DLPX_ERR_LINK_CELLS, "too few link cells");
    }
    const int ncell = c->nc[0] * c->nc[1] * c->nc[2];
    // subcells need 6 distinct subcells per axis around a link cell
    const bool sub = c->list_subcells && c->nc[0] >= 3 && c->nc[1] >= 3 && c->nc[2] >= 3;
    const int nkey = sub ? 8 * ncell : ncell;
    if (ncell + 1 > c->ncell_cap) {
        int r;
        if ((r = grow(c, &c->cell_start, ncell + 1))) return r;
        c->ncell_cap = ncell + 1;
    }
    if (nkey + 1 > c->nkey_cap) {
        int r;
        if ((r = grow(c, &c->cell_count, nkey + 1))) return r;
        if ((r = grow(c, &c->cell_fill, nkey + 1))) return r;
        if ((r = grow(c, &c->key_start, nkey + 1))) return r;
        // the scan zeroes the counts it has read, so they are zero between builds
        DLPX_HIP(c, dev_zero(c->stream, c->cell_count, sizeof(int32_t) * (nkey + 1)));
        c->nkey_cap = nkey + 1;
    }
    c->ncell = ncell;
    CellGrid g = make_grid(c);
    g.sub = sub ? 2 : 1;
    int32_t *kstart = sub ? c->key_start : c->cell_start;
    lc.n = n;
    lc.ncell = ncell;
    lc.nkey = nkey;
    lc.sub = sub;
    const bool res = c->resident;
    lc.wrap = WrapArgs{c->pos,    n,     g, c->cell_of, c->cell_count, kstart, c->cell_fill, nkey,
                       sub ? c->cell_start : nullptr, (unsigned int *)(c->flags + F_WDONE), upd};
    lc.scat = ScatterArgs{c->cell_of, c->cell_fill, c->perm,    n,          upd,
                          c->flags,   c->pos,       c->pos_alt, c->typ,     c->typ_alt,
                          c->gid,     c->gid_alt,   res ? c->vel : nullptr, res ? c->vel_alt : nullptr};
    lc.sort = SortArgs{c->perm,    kstart, nkey,   n,      upd,     c->flags,
                       c->pos_alt, c->pos, c->typ_alt, c->typ, c->gid_alt, c->gid,
                       res ? c->vel_alt : nullptr, res ? c->vel : nullptr, c->slot, c->xref};
    lc.build = BuildArgs{c->pos, g, sub ? c->key_start : c->cell_start, rc_ext * rc_ext, c->cap_lane,
                         c->nbr, c->cnt, c->flags, upd, c->sub_of};
    c->list_coded = sub;
    lc.excl_on = c->excl_total > 0 || c->nfrozen > 1;
    lc.excl = ExclListArgs{n,      c->cap_lane, c->nbr, c->cnt, c->gid,
                           c->excl_total > 0 ? c->excl_off : nullptr, c->excl_ids,
                           c->nfrozen > 1 ? c->frz : nullptr, c->flags, upd};
    return 0;
}

// Bin, sort and permute the atoms of a list chain (everything before the
// list kernel itself).
static void launch_sort(const ListChain &lc, hipStream_t s)
{
    hipLaunchKernelGGL(k_wrap_scan, dim3(ceil_div(lc.n, kWrapThreads)), dim3(kWrapThreads), 0, s, lc.wrap);
    hipLaunchKernelGGL(k_scatter_copy, dim3(ceil_div(lc.n, 256)), dim3(256), 0, s, lc.scat);
    hipLaunchKernelGGL(k_sort_permute, dim3(lc.nkey), dim3(lc.sub ? 64 : 256), 0, s, lc.sort);
}

static void launch_build(const ListChain &lc, hipStream_t s)
{
    if (lc.sub)
        hipLaunchKernelGGL(k_build_list_sub, dim3(xcd_grid(lc.ncell)), dim3(kListThreads), 0, s, lc.build);
    else
        hipLaunchKernelGGL(k_build_list, dim3(xcd_grid(lc.ncell)), dim3(kListThreads), 0, s, lc.build);
    if (lc.excl_on)
        hipLaunchKernelGGL(k_list_exclude, dim3(ceil_div((long)lc.n * kLanesPerAtom, 256)), dim3(256), 0, s, lc.excl);
}

// Bin, sort, permute and build the list from c->pos (current order); the
// host waits for the build and regrows an overflowing list.
int build_list(Ctx *c)
{
    ListChain lc;
    int r;
    if ((r = list_chain(c, nullptr, lc))) return r;
    const int n = c->natms;
    const double rc_ext = c->prm.cutoff + c->prm.padding;
    hipStream_t s = c->stream;
    launch_sort(lc, s);
    constexpr int G = kLanesPerAtom;
    for (int attempt = 0; attempt < 2; attempt++) {
        if (c->cap_lane == 0) {
            // first guess: density-based estimate of the per-atom count (neighbours.F90 max_list sizing)
            const double rho = (double)n / (c->L[0] * c->L[1] * c->L[2]);
            const double est = 4.18879020478639 * rho * rc_ext * rc_ext * rc_ext;  // 4/3 pi r^3 rho
            const int per_atom = c->prm.max_list > 0 ? c->prm.max_list : (int)(1.5 * est) + 64;
            c->cap_lane = (per_atom + G - 1) / G;
            if ((r = grow(c, &c->nbr, (size_t)c->cap_lane * c->cap * G))) return r;
        }
        lc.build.cap_lane = lc.excl.cap_lane = c->cap_lane;
        lc.build.nbr = lc.excl.nbr = c->nbr;
        DLPX_HIP(c, hipMemsetAsync(c->flags, 0, sizeof(int32_t) * 4, s));
        if (c->timing) DLPX_HIP(c, hipEventRecord(c->ev[6], s));
        launch_build(lc, s);
        if (c->timing) DLPX_HIP(c, hipEventRecord(c->ev[7], s));
        DLPX_HIP(c, hipMemcpyAsync(c->h_flags, c->flags, sizeof(int32_t) * 4, hipMemcpyDeviceToHost, s));
        DLPX_HIP(c, hipMemcpyAsync(c->h_flags + 10, c->flags + F_ERR, sizeof(int32_t), hipMemcpyDeviceToHost, s));
        DLPX_HIP(c, hipStreamSynchronize(s));
        if (c->h_flags[10] != 0) {
            c->have_list = false;
            return fail(c, DLPX_ERR_STATE, "internal error: the list build met an atom index out of range");
        }
        if (c->timing) {
            float ms;
            DLPX_HIP(c, hipEventElapsedTime(&ms, c->ev[6], c->ev[7]));
            c->stage_ms[5] += ms;
        }
        if (c->h_flags[F_OVF] == 0) break;  // no lane overflowed
        if (c->prm.max_list > 0) {
            c->have_list = false;
            return fail(c, DLPX_ERR_LIST_OVERFLOW,
                        "neighbour list overflow: an atom has " + std::to_string(c->h_flags[F_IBIG]) +
                            " neighbours, max_list is " + std::to_string(c->prm.max_list));
        }
        // grow with margin and rebuild
        c->cap_lane = c->h_flags[F_OVF] + c->h_flags[F_OVF] / 4 + 8;
        if ((r = grow(c, &c->nbr, (size_t)c->cap_lane * c->cap * G))) return r;
    }
    if (c->h_flags[F_OVF] != 0) {
        c->have_list = false;
        return fail(c, DLPX_ERR_LIST_OVERFLOW, "neighbour list overflow after regrowth");
    }
    c->have_list = true;
    return 0;
}

// Batched list builds of an ensemble, each gated by its context's device
// rebuild decision (F_UPD).  One launch per kernel of the chain over every
// system; the link-cell and subcell list kernels each take the systems that
// use them.
int list_prepare(const std::vector<Ctx *> &cs, ListBatch &lb)
{
    std::vector<WrapArgs> wrap;
    std::vector<ScatterArgs> scat;
    std::vector<SortArgs> sort;
    std::vector<BuildArgs> bsub, blink;
    std::vector<ExclListArgs> excl;
    std::vector<int> nb_wrap, nb_atoms, nb_keys, nb_sub, nb_link, nb_excl;
    int r;
    for (Ctx *c : cs) {
        ListChain lc;
        if ((r = list_chain(c, c->flags + F_UPD, lc))) return r;
        if (c->cap_lane == 0 || !c->nbr) return fail(c, DLPX_ERR_STATE, "no list to rebuild on the device");
        wrap.push_back(lc.wrap);
        scat.push_back(lc.scat);
        sort.push_back(lc.sort);
        excl.push_back(lc.excl);
        nb_wrap.push_back(ceil_div(lc.n, kWrapThreads));
        nb_atoms.push_back(ceil_div(lc.n, 256));
        nb_keys.push_back(lc.nkey);
        nb_sub.push_back(lc.sub ? lc.ncell : 0);
        nb_link.push_back(lc.sub ? 0 : lc.ncell);
        nb_excl.push_back(lc.excl_on ? ceil_div((long)lc.n * kLanesPerAtom, 256) : 0);
        bsub.push_back(lc.build);
        blink.push_back(lc.build);
    }
    Ctx *c0 = cs[0];
    if ((r = batch_upload(c0, lb.wrap, wrap.data(), sizeof(WrapArgs), nb_wrap, false)) ||
        (r = batch_upload(c0, lb.scat, scat.data(), sizeof(ScatterArgs), nb_atoms, false)) ||
        (r = batch_upload(c0, lb.sort, sort.data(), sizeof(SortArgs), nb_keys, false)) ||
        (r = batch_upload(c0, lb.bsub, bsub.data(), sizeof(BuildArgs), nb_sub, true)) ||
        (r = batch_upload(c0, lb.blink, blink.data(), sizeof(BuildArgs), nb_link, true)) ||
        (r = batch_upload(c0, lb.excl, excl.data(), sizeof(ExclListArgs), nb_excl, false)))
        return r;
    // a batch of one system gates every kernel on its own decision word, read
    // first thing by each workgroup; several systems run the compacted chain
    lb.gate = cs.size() == 1 ? c0->flags + F_UPD : nullptr;
    if (cs.size() > 1) {
        const int nsys = (int)cs.size();
        const std::vector<int> *nbs[kListKernels] = {&nb_wrap, &nb_atoms, &nb_keys, &nb_sub, &nb_link, &nb_excl};
        std::vector<int32_t> nb((size_t)kListKernels * nsys);
        for (int q = 0; q < kListKernels; q++)
            for (int k = 0; k < nsys; k++) nb[(size_t)q * nsys + k] = (*nbs[q])[k];
        std::vector<int32_t *> gates(nsys);
        for (int k = 0; k < nsys; k++) gates[k] = cs[k]->flags + F_UPD;
        const size_t plan_words = (size_t)nsys + 1 + (size_t)kListKernels * (nsys + 1);
        DLPX_HIP(c0, hipMalloc((void **)&lb.plan, sizeof(int32_t) * plan_words));
        DLPX_HIP(c0, hipMalloc((void **)&lb.plan_nb, sizeof(int32_t) * nb.size()));
        DLPX_HIP(c0, hipMalloc((void **)&lb.plan_gates, sizeof(int32_t *) * nsys));
        DLPX_HIP(c0, dev_upload(c0->stream, lb.plan_nb, nb.data(), sizeof(int32_t) * nb.size()));
        DLPX_HIP(c0, dev_upload(c0->stream, lb.plan_gates, gates.data(), sizeof(int32_t *) * nsys));
        // fixed grids of the compacted kernels: enough workgroups to fill the
        // GPU a few times over (the subcell and link-cell builds hold 52 KB of
        // LDS each, three to a CU)
        const int cap[kListKernels] = {2048, 2048, 8192, 1024, 1024, 2048};
        for (int q = 0; q < kListKernels; q++) {
            long tot = 0;
            for (int k = 0; k < nsys; k++) tot += (*nbs[q])[k];
            lb.plan_grid[q] = (int)std::min<long>(tot, cap[q]);
        }
        lb.nsys = nsys;
    }
    return 0;
}

int list_launch(Ctx *c, hipStream_t s, const ListBatch &lb, int part)
{
    // part 0: binning and sort (everything the reciprocal-space stages read);
    // part 1: the list build and its exclusion filter; -1: both
    const bool p0 = part != 1, p1 = part != 0;
    if (lb.plan) {
        if (p0)
            hipLaunchKernelGGL(k_list_plan, dim3(1), dim3(256), 0, s, lb.nsys,
                               (int32_t *const *)lb.plan_gates, (const int32_t *)lb.plan_nb, lb.plan);
#define DLPX_LC(K, BL, ARGS, NT, ROW)                                                                     \
    if (lb.plan_grid[ROW] > 0)                                                                           \
        hipLaunchKernelGGL(K, dim3(lb.plan_grid[ROW]), dim3(NT), 0, s, (const ARGS *)(BL).args,          \
                           (const int32_t *)lb.plan, lb.nsys, ROW)
        if (p0) {
            DLPX_LC(k_wrap_scan_cbatch, lb.wrap, WrapArgs, kWrapThreads, 0);
            DLPX_LC(k_scatter_copy_cbatch, lb.scat, ScatterArgs, 256, 1);
            DLPX_LC(k_sort_permute_cbatch, lb.sort, SortArgs, 256, 2);
        }
        if (p1) {
            DLPX_LC(k_build_list_sub_cbatch, lb.bsub, BuildArgs, kListThreads, 3);
            DLPX_LC(k_build_list_cbatch, lb.blink, BuildArgs, kListThreads, 4);
            DLPX_LC(k_list_exclude_cbatch, lb.excl, ExclListArgs, 256, 5);
        }
#undef DLPX_LC
        DLPX_HIP(c, hipGetLastError());
        return 0;
    }
#define DLPX_LB(K, BL, ARGS, NT)                                                                          \
    if ((BL).total > 0)                                                                                  \
        hipLaunchKernelGGL(K, dim3((BL).total), dim3(NT), 0, s, (const ARGS *)(BL).args, (BL).boff, (BL).nsys, \
                           (const int32_t *)lb.gate)
    if (p0) {
        DLPX_LB(k_wrap_scan_batch, lb.wrap, WrapArgs, kWrapThreads);
        DLPX_LB(k_scatter_copy_batch, lb.scat, ScatterArgs, 256);
        DLPX_LB(k_sort_permute_batch, lb.sort, SortArgs, 256);
    }
    if (p1) {
        DLPX_LB(k_build_list_sub_batch, lb.bsub, BuildArgs, kListThreads);
        DLPX_LB(k_build_list_batch, lb.blink, BuildArgs, kListThreads);
        DLPX_LB(k_list_exclude_batch, lb.excl, ExclListArgs, 256);
    }
#undef DLPX_LB
    DLPX_HIP(c, hipGetLastError());
    return 0;
}

void list_batch_free(ListBatch &lb)
{
    for (BatchLaunch *b : {&lb.wrap, &lb.scat, &lb.sort, &lb.bsub, &lb.blink, &lb.excl}) batch_free(*b);
    for (void *p : {(void *)lb.plan, (void *)lb.plan_nb, (void *)lb.plan_gates})
        if (p) (void)hipFree(p);
    lb.plan = lb.plan_nb = nullptr;
    lb.plan_gates = nullptr;
    lb.gate = nullptr;
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

int count_pairs_dev(Ctx *c, double r, int64_t *npairs)
{
    unsigned long long *d = nullptr, h = 0;
    DLPX_HIP(c, hipMalloc((void **)&d, sizeof(*d)));
    hipError_t e = hipMemsetAsync(d, 0, sizeof(*d), c->stream);
    if (e == hipSuccess) {
        const CellGrid g = make_grid(c);
        hipLaunchKernelGGL(k_count_pairs, dim3(ceil_div((long)c->natms * kLanesPerAtom, 256)), dim3(256), 0, c->stream,
                           c->pos, c->nbr, c->cnt, c->natms, c->cap_lane, r * r, g, d);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(&h, d, sizeof(h), hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    (void)hipFree(d);
    if (e != hipSuccess) return fail(c, DLPX_ERR_HIP + (int)e, std::string("count_pairs: ") + hipGetErrorString(e));
    *npairs = (int64_t)(h / 2);
    return 0;
}

}  // namespace dlpx
