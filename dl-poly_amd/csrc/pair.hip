// Real-space two-body forces on the Verlet list: tabulated vdW
// (vdw.F90:1790-2024) and erfc-damped Coulomb (ewald_spole.F90:58-242),
// both by DL_POLY's 3-point interpolation of its own tables.
//
// G lanes share one atom's (full) neighbour list; the force on the atom is
// summed in FP64 across the lanes, energies/virial/stress are halved (each
// pair is met from both ends) and reduced per block.  The tables live in LDS
// as {force, energy} pairs per sample so one lookup is three 16-byte (FP64)
// or 8-byte (FP32) LDS reads.
#include "common.h"
#include "device_util.h"

namespace dlpx {

template <typename R> struct R2T;
template <> struct R2T<double> { using type = double2; };
template <> struct R2T<float> { using type = float2; };

struct PairParams {
    int ntypes, nkeys, mg1, ns1;       // mg1 = max_grid+1, ns1 = nsamples+1
    double L[3], invL[3];
    double rmax2;                      // max(rc, rvdw)^2
    double rc, rvdw;
    double rdr_v, rdr_e;               // reciprocal table spacings
    double scaling;                    // r4pie0/eps
    double r_rvdw;                     // 1/rvdw (force-shifting)
    int electro;
    int cap_lane;                      // list capacity per lane
};

// 3-point Newton interpolation of DL_POLY (numerics.F90:250-290) on the
// .x (force) and .y (energy) members of consecutive samples.
// r and 1/r^2 from r^2: FP64 with IEEE square root and division; FP32 from
// the hardware reciprocal square root refined by one Newton step.
__device__ __forceinline__ void r_rinv2(double rsq, double &r, double &rinv2)
{
    r = sqrt(rsq);
    const double ri = 1.0 / r;
    rinv2 = ri * ri;
}
__device__ __forceinline__ void r_rinv2(float rsq, float &r, float &rinv2)
{
    const float ri0 = __builtin_amdgcn_rsqf(rsq);
    // one Newton step on 1/sqrt(rsq): the hardware estimate's error leans one
    // way, which would bias every table lookup (r) and every pair's force
    // and virial (1/r^2) alike
    const float u = __builtin_fmaf(-(rsq * ri0), ri0, 1.0f);
    const float ri = __builtin_fmaf(0.5f * ri0, u, ri0);
    r = rsq * ri;
    rinv2 = ri * ri;
}

// The table coordinate r/dr: FP64 as the reference forms it (r * rdr);
// FP32 with the reciprocal spacing split into a float and its rounding
// remainder (rdr_lo), so the coordinate carries no systematic offset from
// the float rounding of rdr, which would move every lookup of a table the
// same way and bias the sums of steep forms (about 3e-7 of a ZBL sum).
template <typename R>
__device__ __forceinline__ R table_coord(R r, R rdr, R rdr_lo)
{
    if constexpr (sizeof(R) == 4)
        return __builtin_fmaf(r, rdr, r * rdr_lo);
    else
        return r * rdr;
}

template <typename R, typename R2>
__device__ __forceinline__ void interp2(const R2 *tab, R r, R rdr, R rdr_lo, R &f, R &e)
{
    const R s = table_coord(r, rdr, rdr_lo);
    const int l = (int)s;
    const R ppp = s - (R)l;
    const R2 a = tab[l], b = tab[l + 1], c = tab[l + 2];
    R f0 = a.x;
    if (l == 0) f0 = f0 * r;
    R t1 = f0 + (b.x - f0) * ppp;
    R t2 = b.x + (c.x - b.x) * (ppp - (R)1);
    f = t1 + (t2 - t1) * ppp * (R)0.5;
    R e0 = a.y;
    if (l == 0) e0 = e0 * r;
    t1 = e0 + (b.y - e0) * ppp;
    t2 = b.y + (c.y - b.y) * (ppp - (R)1);
    e = t1 + (t2 - t1) * ppp * (R)0.5;
}

// vdw_method direct: U(r) and gamma = -r dU/dr of the generated forms
// (two_body_potentials.F90, keys of vdw.F90:64-113), in R.
template <typename R>
__device__ __forceinline__ R rpow(R x, int n)
{
    R y = (R)1;
    for (int k = 0; k < n; k++) y *= x;
    return y;
}
// pieces of the composite forms (kk = Z1 Z2 r4pie0 and a = inverse
// screening length prepared on the host)
template <typename R>
__device__ __forceinline__ void zbl_form(R kk, R a, R r, R &e, R &g)
{
    const R b[4] = {(R)0.18175, (R)0.50986, (R)0.28022, (R)0.02817};
    const R c[4] = {(R)3.1998, (R)0.94229, (R)0.40290, (R)0.20162};
    const R x = r * a;
    R s = (R)0, sc = (R)0;
#pragma unroll
    for (int i = 0; i < 4; i++) {
        const R t = b[i] * exp(-x * c[i]);
        s += t;
        sc += c[i] * t;
    }
    e = kk * s * ((R)1 / r);
    g = e + a * kk * sc;
}
template <typename R>
__device__ __forceinline__ void fermi_form(R rm, R ic, R r, R &f, R &df)
{
    const R c = (R)1 / ic;
    const R t = exp(-fabs(r - rm) * c) * (R)0.5;
    f = r < rm ? (R)1 - t : t;
    df = r * c * t;
}
template <typename R>
__device__ __forceinline__ void mdf_form(R ri, R rc, R r, R &m, R &dm)
{
    m = r < ri ? (R)1 : (R)0;
    dm = (R)0;
    if (r >= ri && r <= rc) {
        const R w = rc - ri, rci = w * w * w * w * w, u = rc - r;
        m = u * u * u * ((R)10 * ri * ri - (R)5 * rc * ri - (R)15 * r * ri + rc * rc + (R)3 * r * rc + (R)6 * r * r) / rci;
        dm = (R)30 * r * (r - rc) * (r - rc) * (r - ri) * (r - ri) / rci;
    }
}
template <typename R>
__device__ __forceinline__ void morse_form(const R *p, R r, R &e, R &g)
{
    const R t = exp(-p[2] * (r - p[1]));
    e = p[0] * (((R)1 - t) * ((R)1 - t) - (R)1);
    g = (R)-2 * r * p[0] * p[2] * ((R)1 - t) * t;
}
template <typename R>
__device__ __forceinline__ void buck_form(const R *p, R r, R &e, R &g)
{
    const R b = r / p[1], t1 = p[0] * exp(-b), t2 = -p[2] / rpow(r, 6);
    e = t1 + t2;
    g = t1 * b + (R)6 * t2;
}

template <typename R>
__device__ __forceinline__ void vdw_form(int kind, const R *p, R r, R &e, R &g)
{
    e = (R)0;
    g = (R)0;
    switch (kind) {
        case 15:  // ZBL (:716-741)
            zbl_form(p[0], p[1], r, e, g);
            break;
        case 16:    // ZBL switched to Morse (:787-808)
        case 17: {  // ZBL switched to Buckingham (:826-847)
            R z, dz, f, df, m, dm;
            zbl_form(p[0], p[1], r, z, dz);
            fermi_form(p[2], p[3], r, f, df);
            if (kind == 16)
                morse_form(p + 4, r, m, dm);
            else
                buck_form(p + 4, r, m, dm);
            e = f * z + ((R)1 - f) * m;
            g = f * dz + df * z + ((R)1 - f) * dm - df * m;
            break;
        }
        case 18:    // LJ x MDF (:959-976)
        case 19:    // Buckingham x MDF (:991-1009)
        case 20: {  // 12-6 x MDF (:1025-1042)
            R u, du, m, dm;
            if (kind == 18) {
                const R s = p[1] / r, s2 = s * s, s6 = s2 * s2 * s2;
                u = (R)4 * p[0] * s6 * (s6 - (R)1);
                du = (R)24 * p[0] * s6 * ((R)2 * s6 - (R)1);
                mdf_form(p[2], p[3], r, m, dm);
            } else if (kind == 19) {
                buck_form(p, r, u, du);
                mdf_form(p[3], p[4], r, m, dm);
            } else {
                const R r6 = rpow((R)1 / r, 6);
                u = (p[0] * r6 - p[1]) * r6;
                du = (R)6 * r6 * ((R)2 * p[0] * r6 - p[1]);
                mdf_form(p[2], p[3], r, m, dm);
            }
            e = u * m;
            g = du * m + dm * u;
            break;
        }
        case 21:  // LJF, zero beyond rc (:924-945)
            if (r * r <= p[2]) {
                const R ir = (R)1 / (r * r), st = p[1] * ir, rct = p[2] * ir, x = p[0] * (rct - (R)1) * (rct - (R)1);
                e = x * (st - (R)1);
                g = (R)4 * p[0] * rct * (rct - (R)1) * (st - (R)1) + (R)2 * x * st;
            }
            break;
        case 24: {  // Stillinger-Weber two-body part, zero beyond aa sig (:1058-1082)
            const R rs = p[6] * p[3];
            if (r < rs) {
                const R ex = p[3] / (r - rs), pr = p[3] / r, c = p[2] * pow(pr, p[4]), prq = pow(pr, p[5]);
                const R xe = exp(ex), t = p[1] * p[0] * (c - prq) * xe;
                e = t;
                g = p[1] * p[0] * (p[4] * c - p[5] * prq) * xe + t * r * ex / (r - rs);
            }
            break;
        }
        case 1: {  // 12-6 (:307-317)
            const R r6 = rpow((R)1 / r, 6);
            e = (p[0] * r6 - p[1]) * r6;
            g = (R)6 * r6 * ((R)2 * p[0] * r6 - p[1]);
            break;
        }
        case 2: {  // Lennard-Jones (:260-270)
            const R s = p[1] / r, s2 = s * s, s6 = s2 * s2 * s2;
            e = (R)4 * p[0] * s6 * (s6 - (R)1);
            g = (R)24 * p[0] * s6 * ((R)2 * s6 - (R)1);
            break;
        }
        case 3: {  // n-m, integer powers (:330-345)
            const R a = p[3] / r, b = (R)1 / (p[1] - p[2]);
            const R rn = rpow(a, (int)p[1]), rm = rpow(a, (int)p[2]);
            e = p[0] * (p[2] * rn - p[1] * rm) * b;
            g = p[0] * p[2] * p[1] * (rn - rm) * b;
            break;
        }
        case 4:  // Buckingham (:471-485)
            buck_form(p, r, e, g);
            break;
        case 5: {  // Born-Huggins-Meyer (:499-514)
            const R ri2 = (R)1 / (r * r);
            const R t1 = p[0] * exp(p[1] * (p[2] - r)), t2 = -p[3] * (ri2 * ri2 * ri2),
                    t3 = -p[4] * (ri2 * ri2 * ri2 * ri2);
            e = t1 + t2 + t3;
            g = t1 * r * p[1] + (R)6 * t2 + (R)8 * t3;
            break;
        }
        case 6: {  // 12-10 (:530-546)
            const R ri2 = (R)1 / (r * r), f12 = p[0] * rpow(ri2, 6), f10 = -p[1] * rpow(ri2, 5);
            e = f12 + f10;
            g = (R)12 * f12 + (R)10 * f10;
            break;
        }
        case 8:  // Morse (:416-428)
            morse_form(p, r, e, g);
            break;
        case 9:  // WCA, zero beyond cut = 2^(1/6) sig + d (:559-576)
            if (r < p[3] || fabs(r - p[2]) < (R)1.0e-10) {
                const R s6 = rpow(p[1] / (r - p[2]), 6);
                e = (R)4 * p[0] * s6 * (s6 - (R)1) + p[0];
                g = (R)24 * p[0] * s6 * ((R)2 * s6 - (R)1) * r / (r - p[2]);
            }
            break;
        case 10:  // DPD, zero beyond rc (:591-610)
            if (r < p[1]) {
                const R x = r / p[1], h = (R)0.5 * p[0] * p[1] * ((R)1 - x);
                e = h * ((R)1 - x);
                g = (R)2 * h * x;
            }
            break;
        case 12: {  // Lennard-Jones cohesive (:283-293)
            const R s6 = rpow(p[1] / r, 6);
            e = (R)4 * p[0] * s6 * (s6 - p[2]);
            g = (R)24 * p[0] * s6 * ((R)2 * s6 - p[2]);
            break;
        }
        case 13: {  // Morse + c/r^12 (:442-456)
            const R t = exp(-p[2] * (r - p[1])), c12 = p[3] / rpow(r, 12);
            e = p[0] * t * (t - (R)2) + c12;
            g = (R)-2 * r * p[0] * p[2] * ((R)1 - t) * t + (R)12 * c12;
            break;
        }
        case 14: {  // Rydberg (:689-702)
            const R x = r / p[2], t = exp(-x);
            e = (p[0] + p[1] * r) * t;
            g = x * t * (p[0] - p[1] * p[2] + p[1] * r);
            break;
        }
        case 7:  // n-m shifted and force-corrected at r_trunc (:360-400)
            if (r <= p[4]) {
                const R n = p[1], m = p[2], ni = rint(n), mi = rint(m), t = n - m, b = (R)1 / t;
                const R c = p[4] / p[3], ci = p[3] / p[4];
                const R beta = c * pow((pow(c, mi + (R)1) - (R)1) / (pow(c, ni + (R)1) - (R)1), b);
                const R alpha = -t / (m * pow(beta, ni) * ((R)1 + (n * ci - n - (R)1) * pow(ci, ni)) -
                                      n * pow(beta, mi) * ((R)1 + (m * ci - m - (R)1) * pow(ci, mi)));
                const R e1 = p[0] * alpha, a = p[3] / r, bn = pow(beta * ci, ni) - pow(beta * ci, mi);
                e = e1 * (m * pow(beta, ni) * (pow(a, ni) - pow(ci, ni)) - n * pow(beta, mi) * (pow(a, mi) - pow(ci, mi)) +
                          n * m * ((r / p[4] - (R)1) * bn)) * b;
                g = e1 * m * n * (pow(beta, ni) * pow(a, ni) - pow(beta, mi) * pow(a, mi) - r / p[4] * bn) * b;
            }
            break;
        case 11: {  // AMOEBA 14-7 (:657-676)
            const R rho = r / p[1], t1 = (R)1 / ((R)0.07 + rho), t2 = (R)1 / ((R)0.12 + rpow(rho, 7));
            const R t3 = p[0] * rpow((R)1.07 * t1, 7), t = t3 * ((R)1.12 * t2 - (R)2);
            e = t;
            g = (R)7 * (t1 * t + (R)1.12 * t3 * t2 * t2 * rpow(rho, 6)) * rho;
            break;
        }
        case 22: {  // Sanderson (:865-879)
            const R x = (r - p[1]) / p[2], t = p[0] * exp(-x * x);
            e = -t;
            g = (R)-2 * (r - p[1]) * r * t / (p[2] * p[2]);
            break;
        }
        case 23:  // nDPD, zero beyond rc (:623-642)
            if (r < p[3]) {
                const R x = r / p[3], t1 = p[0] * p[3] * ((R)1 - x), t0 = p[1] * pow((R)1 - x, p[2] - (R)1);
                e = t1 * ((R)1 - x) * (t0 / (p[2] + (R)1) - (R)0.5);
                g = t1 * x * (t0 - (R)1);
            }
            break;
        default:
            break;
    }
}

// 512-thread workgroups: the LDS tables (about 77 KB in FP64 for three vdW
// keys plus erfc) are loaded once per workgroup, so larger groups put more
// wavefronts per CU behind one copy to hide the latency of the neighbour
// position gathers.
constexpr int kPairThreads = 512;
// the combined-table kernel: 1024 threads, so the two workgroups its LDS rows
// leave room for on a CU hold 8 wavefronts per SIMD
constexpr int kCombThreads = 1024;

// GP lanes serve one atom (a multiple of the list's kLanesPerAtom lanes):
// pair lane p walks list lane p % kLanesPerAtom, entries c = p / kLanesPerAtom,
// c + GP / kLanesPerAtom, ...  More lanes per atom give more wavefronts for
// the same atoms (latency hiding on small systems) at the same total work.
//
// PP (per-particle collection, stats%collect_pp): each atom also gets half of
// every pair's energy and of r (x) f (vdw.F90:1987-1997, 1750-1760;
// ewald_spole.F90:205-213), written to ppo[i*kPPStride + 0..6].  As in the
// reference's tabulated vdW loop, whose pair energy is evaluated only on the
// side that passes the halo tie-break (vdw.F90:1946-1966), the vdW energy of
// a pair across the cell boundary (nonzero minimum-image shift) goes only to
// the atom with the smaller global id (gid, the host index).
//
// GT: the vdW tables are read from global memory (through the caches) instead
// of LDS, for pair-key sets whose tables do not fit next to the erfc table.
template <typename R>
struct PairArgs {
    const double4 *pos;
    const int32_t *typ;
    int n;
    const int32_t *nbr, *cnt;
    PairParams pp;
    const typename R2T<R>::type *vdw_g, *erfc_g;
    const int32_t *keytab_g;
    const R *direct_g;
    double4 *frc;
    double *red;
    const int32_t *gid;
    double *ppo;
    const float4 *comb_g;        // combined rows (mixed precision, see k_pair_comb)
    const int32_t *comb_row_g;   // ntypes*ntypes -> combined row
    const float4 *prel;          // relative positions (k_pair_comb<REL>)
    float hs[3], hs_lo[3];       // subcell edge: its FP32 rounding and the remainder
};
template <typename R, int GP, int NT, bool FS, bool DIRECT, bool PP = false, bool GT = false>
__device__ __forceinline__ void pair_body(const PairArgs<R> &a, int bid)
{
    using R2 = typename R2T<R>::type;
    const double4 *__restrict__ pos = a.pos;
    const int32_t *__restrict__ typ = a.typ;
    const int n = a.n;
    const int32_t *__restrict__ nbr = a.nbr;
    const int32_t *__restrict__ cnt = a.cnt;
    const PairParams &pp = a.pp;
    const R2 *__restrict__ vdw_g = a.vdw_g;
    const R2 *__restrict__ erfc_g = a.erfc_g;
    const int32_t *__restrict__ keytab_g = a.keytab_g;
    const R *__restrict__ direct_g = a.direct_g;
    double4 *__restrict__ frc = a.frc;
    double *__restrict__ red = a.red;
    const int32_t *__restrict__ gid = a.gid;
    double *__restrict__ ppo = a.ppo;
    extern __shared__ __align__(16) unsigned char smem[];
    // LDS: vdW tables (none in direct mode or with GT), erfc table, key table,
    // direct forms (8 per table row), reduction scratch
    R2 *tvl = (R2 *)smem;
    const R2 *tv = GT ? vdw_g : tvl;
    R2 *te = tvl + ((DIRECT || GT) ? 0 : pp.nkeys * pp.mg1);
    int32_t *kt = (int32_t *)(te + (pp.electro ? pp.ns1 : 0));
    R *dv = (R *)(kt + ((pp.ntypes * pp.ntypes + 1) & ~1));
    double *lred = (double *)(dv + (DIRECT ? kDirectStride * pp.nkeys : 0));

    if (DIRECT)
        for (int k = threadIdx.x; k < kDirectStride * pp.nkeys; k += blockDim.x) dv[k] = direct_g[k];
    else if (!GT)
        for (int k = threadIdx.x; k < pp.nkeys * pp.mg1; k += blockDim.x) tvl[k] = vdw_g[k];
    if (pp.electro)
        for (int k = threadIdx.x; k < pp.ns1; k += blockDim.x) te[k] = erfc_g[k];
    for (int k = threadIdx.x; k < pp.ntypes * pp.ntypes; k += blockDim.x) kt[k] = keytab_g[k];
    __syncthreads();

    constexpr int G = kLanesPerAtom, CS = GP / G;
    static_assert(GP % G == 0 && GP <= 64, "GP must be a multiple of the list lanes");
    const int tid = bid * blockDim.x + threadIdx.x;
    const int i = tid / GP, p = tid % GP, lane = p % G, c0 = p / G;
    double fx = 0.0, fy = 0.0, fz = 0.0;
    // engvdw, virvdw, engrl, (unused: the Coulomb virial is -tr(stress) - virvdw,
    // formed after the reduction), s11..s33
    // per-lane partial sums over this lane's ~m/CS pairs in R (FP32 in mixed
    // precision: one add per term instead of a conversion and an FP64 add),
    // widened to FP64 before any sum across lanes; forces accumulate in FP64
    R pacc[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    R ppe = 0;  // PP: this atom's energy share (before the halving)
    if (i < n) {
        const double4 pi = pos[i];
        const int ti = typ[i];
        const R qi = (R)(pi.w * pp.scaling);
        const bool qi_on = fabs(pi.w * pp.scaling) >= kZeroPlus && pp.electro;
        const int m = min(cnt[(size_t)i * G + lane], pp.cap_lane);
        const R rvdw = (R)pp.rvdw, rc = (R)pp.rc, rdr_v = (R)pp.rdr_v, rdr_e = (R)pp.rdr_e;
        const R rdr_v_lo = (R)(pp.rdr_v - (double)rdr_v), rdr_e_lo = (R)(pp.rdr_e - (double)rdr_e);
        const int32_t *ktrow = kt + ti * pp.ntypes;
        const int32_t *lst = nbr + list_slot(i, 0, lane, pp.cap_lane);  // entry c at lst[c * G]
        const R rmax2 = (R)pp.rmax2;
        // software pipeline: the neighbour index is loaded two iterations
        // ahead and its position and type one iteration ahead, so neither
        // load's latency sits in front of the pair arithmetic
        // The prefetches are unconditional, from the lane's slot clamped to its
        // last entry (always a real neighbour): a conditional load makes the
        // compiler wait for it where the branches join, right after issuing
        // it.  The entry is kept raw and masked where it is used, for the same
        // reason.  A lane without entries skips the loop and its prologue.
        const int mm = m > 0 ? m - 1 : 0;
        int32_t jn = m > 0 ? lst[(size_t)min(c0, mm) * G] : i;
        double4 pn = pos[list_j(jn)];
        int tn = typ[list_j(jn)];
        int gn = PP ? gid[list_j(jn)] : 0;
        const int gi = PP ? gid[i] : 0;
        jn = m > 0 ? lst[(size_t)min(c0 + CS, mm) * G] : i;
        for (int c = c0; c < m; c += CS) {
            const double4 pj = pn;
            const int tj = tn;
            const int gj = gn;
            pn = pos[list_j(jn)];
            tn = typ[list_j(jn)];
            if (PP) gn = gid[list_j(jn)];
            jn = lst[(size_t)min(c + 2 * CS, mm) * G];
            const double dxd = min_image_1(pi.x, pj.x, pp.L[0], pp.invL[0]);
            const double dyd = min_image_1(pi.y, pj.y, pp.L[1], pp.invL[1]);
            const double dzd = min_image_1(pi.z, pj.z, pp.L[2], pp.invL[2]);
            const R dx = (R)dxd, dy = (R)dyd, dz = (R)dzd;
            const R rsq = dx * dx + dy * dy + dz * dz;
            if (rsq >= rmax2) continue;
            R rrr, rinv2;
            r_rinv2(rsq, rrr, rinv2);
            R gam = (R)0;
            const int key = ktrow[tj];
            if (key >= 0 && rrr < rvdw) {
                R gf, ge;
                R g, ev;
                if (DIRECT) {
                    // vdw.F90:1696-1699
                    const R *dk = dv + kDirectStride * key;
                    const R afs = dk[kDirectStride - 2], bfs = dk[kDirectStride - 1];
                    vdw_form<R>((int)dk[0], dk + 1, rrr, ge, gf);
                    ev = ge + afs * rrr + bfs;
                    g = gf * rinv2 - afs * (rrr * rinv2);
                } else {
                    interp2<R, R2>(tv + key * pp.mg1, rrr, rdr_v, rdr_v_lo, gf, ge);
                    g = gf * rinv2;
                    ev = ge;
                }
                if (FS && !DIRECT) {
                    // vdw.F90:1922,1961-1964 with the samples at r = rvdw (index max_grid-4)
                    const R2 cut = tv[key * pp.mg1 + pp.mg1 - 5];
                    const R r_rvdw = (R)pp.r_rvdw;
                    g = g - cut.x * ((rrr * rinv2) * r_rvdw);
                    // r/rvdw - 1 with 1/rvdw split like the table coordinate
                    const R rs1 = sizeof(R) == 4 ? fma(rrr, (R)(pp.r_rvdw - (double)r_rvdw), fma(rrr, r_rvdw, (R)-1))
                                                 : rrr * r_rvdw - (R)1;
                    ev = ev + cut.x * rs1 - cut.y;
                }
                gam = g;
                pacc[0] += ev;
                pacc[1] -= g * (rrr * rrr);
                if (PP) {
                    const bool cross = dxd != pi.x - pj.x || dyd != pi.y - pj.y || dzd != pi.z - pj.z;
                    if (DIRECT || !cross || gi < gj) ppe += ev;
                }
            }
            // no |q_j| > Tiny test: such a q_j contributes exact (or subnormal) zeros
            if (qi_on && rrr < rc) {
                R df, de;
                interp2<R, R2>(te, rrr, rdr_e, rdr_e_lo, df, de);
                const R prefac = qi * (R)pj.w;
                const R eg = prefac * df;
                gam += eg;
                pacc[2] += prefac * de;
                if (PP) ppe += prefac * de;
            }
            const R fxp = gam * dx, fyp = gam * dy, fzp = gam * dz;
            fx += (double)fxp;
            fy += (double)fyp;
            fz += (double)fzp;
            pacc[4] += dx * fxp;
            pacc[5] += dx * fyp;
            pacc[6] += dx * fzp;
            pacc[7] += dy * fyp;
            pacc[8] += dy * fzp;
            pacc[9] += dz * fzp;
        }
    }
    fx = group_sum<GP>(fx);
    fy = group_sum<GP>(fy);
    fz = group_sum<GP>(fz);
    if (i < n && p == 0) frc[i] = make_double4(fx, fy, fz, 0.0);
    if (PP) {
        double v[7] = {(double)ppe, (double)pacc[4], (double)pacc[5], (double)pacc[6], (double)pacc[7],
                       (double)pacc[8], (double)pacc[9]};
#pragma unroll
        for (int k = 0; k < 7; k++) v[k] = group_sum<GP>(v[k]);
        if (i < n && p == 0)
#pragma unroll
            for (int k = 0; k < 7; k++) ppo[(size_t)i * kPPStride + k] = 0.5 * v[k];
    }
    double acc[10];
#pragma unroll
    for (int k = 0; k < 10; k++) acc[k] = 0.5 * (double)pacc[k];  // every pair is met twice
    __syncthreads();
    block_sum<10>(acc, lred, red + (size_t)bid * kReduceSlots + R_ENGVDW);
}
template <typename R, int GP, int NT, bool FS, bool DIRECT, bool PP = false, bool GT = false>
__global__ void __launch_bounds__(NT) k_pair(PairArgs<R> a)
{
    pair_body<R, GP, NT, FS, DIRECT, PP, GT>(a, xcd_block(blockIdx.x, gridDim.x));
}
template <typename R, int GP, int NT, bool FS, bool DIRECT, bool PP = false, bool GT = false>
__global__ void __launch_bounds__(NT) k_pair_batch(const PairArgs<R> *__restrict__ av,
                                                    const int32_t *__restrict__ boff, int nsys)
{
    BatchPos bp;
    if (!batch_pos(boff, nsys, bp)) return;
    const PairArgs<R> &a = av[bp.s];
    pair_body<R, GP, NT, FS, DIRECT, PP, GT>(a, bp.b);
}

// The mixed-precision pair kernel for the common case of charges fixed per
// atom type, tabulated vdW without force-shifting, and vdW and erfc tables on
// one sample grid (rc = rvdw): one row per type pair holds, per sample r_l,
//   x: H = Fv + q_a q_b r_l^2 Fe, the pair's force factor times r^2
//      (f = H d / r^2, as the vdW table's Fv = r dU/dr),
//   y: Fv (the vdW virial),
//   z: Ev, the vdW energy,  w: Ec = q_a q_b erfc(alpha r)/r (scaled),
// so a pair costs one table position and three 16-byte LDS reads instead of
// two lookups.  The 3-point Newton interpolation of numerics.F90:250-290 is
// written as a + p (b - a) + p (p - 1)/2 (a - 2b + c) (the same polynomial).
// The vdW part of H interpolates exactly as the vdW table does; the Coulomb
// part interpolates the smooth r^2 Fe instead of Fe, which moves a force by
// the erfc table's interpolation error (~1e-8), inside the mixed bar.
//
// REL: positions come from prel, each atom's position relative to its
// subcell at the list build (FP32) with its type in w, and a list entry's
// code gives the subcell offset d of j from i, so x_i - x_j = (xr_i - xr_j)
// - d h: one 16-byte gather per pair and no FP64 minimum image.  d h is
// applied as d h_hi + d h_lo (h_hi the FP32 rounding of the subcell edge,
// h_lo its remainder): a rounded d h alone would move every pair of one code
// by the same amount, which sums into the virial instead of averaging out.
// What remains is the unbiased FP32 rounding of the relative positions and
// of the sums (below 5e-7 A each for subcells of a few A).
template <int GP, int NT, bool REL>
__device__ __forceinline__ void pair_comb_body(const PairArgs<float> &a, int bid)
{
    const double4 *__restrict__ pos = a.pos;
    const int32_t *__restrict__ typ = a.typ;
    const int n = a.n;
    const int32_t *__restrict__ nbr = a.nbr;
    const int32_t *__restrict__ cnt = a.cnt;
    const PairParams &pp = a.pp;
    double4 *__restrict__ frc = a.frc;
    extern __shared__ __align__(16) unsigned char smem[];
    // LDS: (REL) the 125 subcell offsets -d h_hi and -d h_lo, each padded to
    // 128; the combined rows; the row table; reduction scratch
    float4 *offt = (float4 *)smem;
    float4 *tc = offt + (REL ? 256 : 0);
    int32_t *kt = (int32_t *)(tc + (size_t)pp.nkeys * pp.mg1);
    double *lred = (double *)(kt + ((pp.ntypes * pp.ntypes + 1) & ~1));
    for (int k = threadIdx.x; k < pp.nkeys * pp.mg1; k += blockDim.x) tc[k] = a.comb_g[k];
    for (int k = threadIdx.x; k < pp.ntypes * pp.ntypes; k += blockDim.x) kt[k] = a.comb_row_g[k];
    if (REL && threadIdx.x < 125) {
        const int k = threadIdx.x;
        const float d[3] = {(float)(k % 5 - 2), (float)((k / 5) % 5 - 2), (float)(k / 25 - 2)};
        offt[k] = make_float4(-d[0] * a.hs[0], -d[1] * a.hs[1], -d[2] * a.hs[2], 0.0f);  // exact: |d| <= 2
        offt[128 + k] = make_float4(-d[0] * a.hs_lo[0], -d[1] * a.hs_lo[1], -d[2] * a.hs_lo[2], 0.0f);
    }
    __syncthreads();

    constexpr int G = kLanesPerAtom, CS = GP / G;
    static_assert(GP % G == 0 && GP <= 64, "GP must be a multiple of the list lanes");
    const int tid = bid * blockDim.x + threadIdx.x;
    const int i = tid / GP, p = tid % GP, lane = p % G, c0 = p / G;
    double fx = 0.0, fy = 0.0, fz = 0.0;
    // engvdw, virvdw, engrl, (unused), s11..s33 as in pair_body
    float pacc[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    if (i < n) {
        const int m = min(cnt[(size_t)i * G + lane], pp.cap_lane);
        const float rdr = (float)pp.rdr_v, rdr_lo = (float)(pp.rdr_v - (double)rdr);
        const float rmax2 = (float)pp.rmax2;
        const int32_t *lst = nbr + list_slot(i, 0, lane, pp.cap_lane);
        // software pipeline: pair_body's (entry two iterations ahead, position
        // and type one ahead); REL, whose iterations are shorter, one deeper
        // (entry three ahead, relative position two ahead)
        // (prefetches unconditional from clamped slots, entries masked where
        // used: see pair_body)
        const int mm = m > 0 ? m - 1 : 0;
        double4 pi, pn;
        float4 qi, qa, qb;
        int tn = 0, ka = 0, kb = 0;
        int32_t en;
        if constexpr (REL) {
            qi = a.prel[i];
            const int32_t e0 = m > 0 ? lst[(size_t)min(c0, mm) * G] : i;
            const int32_t e1 = m > 0 ? lst[(size_t)min(c0 + CS, mm) * G] : i;
            en = m > 0 ? lst[(size_t)min(c0 + 2 * CS, mm) * G] : i;
            qa = a.prel[list_j(e0)];
            ka = list_code(e0);
            qb = a.prel[list_j(e1)];
            kb = list_code(e1);
        } else {
            pi = pos[i];
            const int32_t e0 = m > 0 ? lst[(size_t)min(c0, mm) * G] : i;
            pn = pos[list_j(e0)];
            tn = typ[list_j(e0)];
            en = m > 0 ? lst[(size_t)min(c0 + CS, mm) * G] : i;
        }
        const int32_t *ktrow = kt + (REL ? __float_as_int(qi.w) : typ[i]) * pp.ntypes;
        // one pair from its separation and j's type; its force goes to the
        // FP32 partial (gx, gy, gz), which the caller adds to the FP64 sums
        // every few pairs
        float gx = 0.0f, gy = 0.0f, gz = 0.0f;
        auto term = [&](const float dx, const float dy, const float dz, const int tj) {
            const float rsq = dx * dx + dy * dy + dz * dz;
            if (rsq >= rmax2) return;
            float rrr, rinv2;
            r_rinv2(rsq, rrr, rinv2);
            const float4 *row = tc + ktrow[tj] * pp.mg1;
            const float s = table_coord(rrr, rdr, rdr_lo);
            const int l = (int)s;
            const float q1 = s - (float)l;
            float4 A = row[l];
            const float4 B = row[l + 1], C = row[l + 2];
            if (__builtin_expect(__any(l == 0), 0)) {
                // numerics.F90:262: the first sample scaled by r (a pair
                // closer than one table step; the wavefront-uniform test
                // keeps the common case free of the select)
                if (l == 0) {
                    A.x *= rrr;
                    A.y *= rrr;
                    A.z *= rrr;
                    A.w *= rrr;
                }
            }
            const float q2 = 0.5f * q1 * (q1 - 1.0f);
            const float gam = rinv2 * __builtin_fmaf(q2, __builtin_fmaf(-2.0f, B.x, A.x + C.x),
                                                     __builtin_fmaf(q1, B.x - A.x, A.x));
            const float fv = __builtin_fmaf(q2, __builtin_fmaf(-2.0f, B.y, A.y + C.y), __builtin_fmaf(q1, B.y - A.y, A.y));
            const float ev = __builtin_fmaf(q2, __builtin_fmaf(-2.0f, B.z, A.z + C.z), __builtin_fmaf(q1, B.z - A.z, A.z));
            const float ec = __builtin_fmaf(q2, __builtin_fmaf(-2.0f, B.w, A.w + C.w), __builtin_fmaf(q1, B.w - A.w, A.w));
            pacc[0] += ev;
            pacc[1] -= fv;
            pacc[2] += ec;
            const float fxp = gam * dx, fyp = gam * dy, fzp = gam * dz;
            gx += fxp;
            gy += fyp;
            gz += fzp;
            pacc[4] += dx * fxp;
            pacc[5] += dx * fyp;
            pacc[6] += dx * fzp;
            pacc[7] += dy * fyp;
            pacc[8] += dy * fzp;
            pacc[9] += dz * fzp;
        };
        if constexpr (REL) {
            // four entries per iteration in two phases of two: phase A works
            // on entries c, c + CS (qa, qb) while it loads the relative
            // positions of c + 2CS, c + 3CS (qc, qd) and the list entries
            // four ahead; phase B works on qc, qd while it loads the next
            // qa, qb.  Every load lands in the register its consumer reads,
            // so the loop carries no copies of values in flight (a copy at the
            // loop's end would wait for the load to arrive there).  All loads
            // are unconditional from clamped slots.
            int32_t ea = en, eb = m > 0 ? lst[(size_t)min(c0 + 3 * CS, mm) * G] : i;
            auto rel_term = [&](const float4 qj, const int code) {
                const float4 of = offt[code], ol = offt[128 + code];
                term(((qi.x - qj.x) + of.x) + ol.x, ((qi.y - qj.y) + of.y) + ol.y, ((qi.z - qj.z) + of.z) + ol.z,
                     __float_as_int(qj.w));
            };
            for (int c = c0; c < m; c += 4 * CS) {
                // list entries first, then the gathers they feed: a wait for
                // an entry (in issue order) then never waits for a gather
                const int32_t ec = lst[(size_t)min(c + 4 * CS, mm) * G], ed = lst[(size_t)min(c + 5 * CS, mm) * G];
                const float4 qc = a.prel[list_j(ea)], qd = a.prel[list_j(eb)];
                const int kc = list_code(ea), kd = list_code(eb);
                rel_term(qa, ka);
                if (c + CS < m) rel_term(qb, kb);
                ea = lst[(size_t)min(c + 6 * CS, mm) * G];
                eb = lst[(size_t)min(c + 7 * CS, mm) * G];
                qa = a.prel[list_j(ec)];
                qb = a.prel[list_j(ed)];
                ka = list_code(ec);
                kb = list_code(ed);
                if (c + 2 * CS < m) rel_term(qc, kc);
                if (c + 3 * CS < m) rel_term(qd, kd);
                // the four pairs' forces into the FP64 sums
                fx += (double)gx;
                fy += (double)gy;
                fz += (double)gz;
                gx = gy = gz = 0.0f;
            }
        } else {
            for (int c = c0; c < m; c += CS) {
                const double4 pj = pn;
                const int tj = tn;
                pn = pos[list_j(en)];
                tn = typ[list_j(en)];
                en = lst[(size_t)min(c + 2 * CS, mm) * G];
                term((float)min_image_1(pi.x, pj.x, pp.L[0], pp.invL[0]),
                     (float)min_image_1(pi.y, pj.y, pp.L[1], pp.invL[1]),
                     (float)min_image_1(pi.z, pj.z, pp.L[2], pp.invL[2]), tj);
                fx += (double)gx;
                fy += (double)gy;
                fz += (double)gz;
                gx = gy = gz = 0.0f;
            }
        }
    }
    fx = group_sum<GP>(fx);
    fy = group_sum<GP>(fy);
    fz = group_sum<GP>(fz);
    if (i < n && p == 0) frc[i] = make_double4(fx, fy, fz, 0.0);
    double acc[10];
#pragma unroll
    for (int k = 0; k < 10; k++) acc[k] = 0.5 * (double)pacc[k];  // every pair is met twice
    __syncthreads();
    block_sum<10>(acc, lred, a.red + (size_t)bid * kReduceSlots + R_ENGVDW);
}
template <int GP, int NT, bool REL>
__global__ void __launch_bounds__(NT) k_pair_comb(PairArgs<float> a)
{
    pair_comb_body<GP, NT, REL>(a, xcd_block(blockIdx.x, gridDim.x));
}
template <int GP, int NT, bool REL>
__global__ void __launch_bounds__(NT) k_pair_comb_batch(const PairArgs<float> *__restrict__ av,
                                                         const int32_t *__restrict__ boff, int nsys)
{
    BatchPos bp;
    if (!batch_pos(boff, nsys, bp)) return;
    pair_comb_body<GP, NT, REL>(av[bp.s], bp.b);
}

// Positions relative to each atom's subcell at the list build (FP32), with
// the type in w, for k_pair_comb<REL>: u = x - o, o = -L/2 + s h the subcell
// origin (the binning of cell_coord), folded to the image nearest the
// subcell (positions need not be folded into the cell between builds).
struct PrelArgs {
    const double4 *pos;
    const int32_t *typ, *sub_of;
    float4 *prel;
    int n;
    double L[3], invL[3], h[3];
};
__device__ __forceinline__ void prel_body(const PrelArgs &a, int bid)
{
    const int i = bid * 256 + threadIdx.x;
    if (i >= a.n) return;
    const double4 p = a.pos[i];
    const int s = a.sub_of[i];
    const int sc[3] = {s & 1023, (s >> 10) & 1023, s >> 20};
    const double x[3] = {p.x, p.y, p.z};
    float u[3];
#pragma unroll
    for (int d = 0; d < 3; d++) {
        double v = x[d] - (sc[d] * a.h[d] - 0.5 * a.L[d]);
        v -= a.L[d] * rint((v - 0.5 * a.h[d]) * a.invL[d]);
        u[d] = (float)v;
    }
    a.prel[i] = make_float4(u[0], u[1], u[2], __int_as_float(a.typ[i]));
}
__global__ void __launch_bounds__(256) k_prel(PrelArgs a) { prel_body(a, blockIdx.x); }
__global__ void __launch_bounds__(256) k_prel_batch(const PrelArgs *__restrict__ av, const int32_t *__restrict__ boff,
                                                    int nsys)
{
    const int lb = blockIdx.x;
    if (lb >= boff[nsys]) return;
    const int s = batch_find(boff, nsys, lb);
    prel_body(av[s], lb - boff[s]);
}

static PrelArgs prel_args(const Ctx *c)
{
    PrelArgs a;
    a.pos = c->pos;
    a.typ = c->typ;
    a.sub_of = c->sub_of;
    a.prel = c->prel;
    a.n = c->natms;
    for (int d = 0; d < 3; d++) {
        a.L[d] = c->L[d];
        a.invL[d] = c->invL[d];
        a.h[d] = c->L[d] / (2.0 * c->nc[d]);  // subcell edge
    }
    return a;
}

// The combined rows (see pair_comb_body) when the context qualifies: mixed
// precision, electrostatics, charges fixed per type, tabulated vdW and erfc
// tables on one grid (equal cutoffs and sample counts).  Built at set-up (the
// list build, resident-run start) from the host tables in FP64, then rounded.
int upload_comb(Ctx *c)
{
    c->comb_dirty = false;
    c->comb_valid = false;
    const int nt = c->prm.ntypes, mg1 = c->max_grid + 1;
    if (c->prm.precision != DLPX_MIXED || !c->prm.electrostatics || !c->q_per_type || c->max_grid != c->nsamples ||
        c->prm.cutoff != c->prm.vdw_cutoff || c->nkeys != nt * (nt + 1) / 2 || (int)c->type_q.size() != nt)
        return 0;
    const int nrows = nt * (nt + 1) / 2;
    const size_t bytes = sizeof(float4) * (size_t)nrows * mg1;
    // the rows must fit the LDS beside the row table and the reduction scratch
    if (bytes + sizeof(float4) * 256 + sizeof(int32_t) * (nt * nt + 1) + sizeof(double) * 10 * (kCombThreads / 64) >
        160 * 1024)
        return 0;
    const double scaling = kR4pie0 / c->prm.epsilon, dr = c->prm.cutoff / (double)(c->max_grid - 4);
    std::vector<float4> rows((size_t)nrows * mg1);
    std::vector<int32_t> rt((size_t)nt * nt);
    for (int a = 1; a <= nt; a++)
        for (int b = 1; b <= a; b++) {
            const int key = a * (a - 1) / 2 + b - 1;  // the reference's pair key of types (a, b)
            rt[(size_t)(a - 1) * nt + (b - 1)] = rt[(size_t)(b - 1) * nt + (a - 1)] = key;
            const double qq = c->type_q[a - 1] * c->type_q[b - 1] * scaling;
            const bool vdw = c->ltp[key] >= 0;  // vdw.F90:1883: an undefined pair has no vdW part
            for (int l = 0; l < mg1; l++) {
                // sample l at r = l dr; erfc samples 1.. are the reference's erfc(1:) (index 0 virtual)
                const double fv = vdw ? c->vdw_frc[(size_t)key * mg1 + l] : 0.0;
                const double ev = vdw ? c->vdw_pot[(size_t)key * mg1 + l] : 0.0;
                const double fe = l > 0 ? c->erfc_d[l - 1] : 0.0, ee = l > 0 ? c->erfc_t[l - 1] : 0.0;
                const double r = l * dr;
                const double g = fv + qq * (r * r) * fe;
                auto fin = [](double x) { return std::isfinite(x) ? (float)std::max(-3.0e38, std::min(3.0e38, x)) : 3.0e38f; };
                float4 v;
                v.x = fin(g);
                v.y = fin(fv);
                v.z = fin(ev);
                v.w = fin(qq * ee);
                rows[(size_t)key * mg1 + l] = v;
            }
        }
    if (c->comb_bytes != bytes) {
        if (c->d_comb) (void)hipFree(c->d_comb);
        c->d_comb = nullptr;
        c->comb_bytes = 0;
        DLPX_HIP(c, hipMalloc(&c->d_comb, bytes));
        c->comb_bytes = bytes;
    }
    if (!c->d_comb_row) DLPX_HIP(c, hipMalloc((void **)&c->d_comb_row, sizeof(int32_t) * (size_t)nt * nt));
    DLPX_HIP(c, dev_upload(c->stream, c->d_comb, rows.data(), bytes));
    DLPX_HIP(c, dev_upload(c->stream, c->d_comb_row, rt.data(), sizeof(int32_t) * rt.size()));
    c->comb_valid = true;
    return 0;
}

// The combined-table kernel runs for this context's evaluations.
static int comb_threads(const Ctx *c) { return c->comb_threads == 512 ? 512 : kCombThreads; }
static bool comb_ok(const Ctx *c)
{
    return c->prm.precision == DLPX_MIXED && c->comb_valid && !c->comb_dirty && !c->vdw_direct &&
           !c->vdw_force_shift && !c->pp_now;
}

// The pair kernel's arguments, dynamic LDS and template variant for one
// context.  Variant: lanes per atom (4, 8, 16), force-shifting, direct
// evaluation, per-particle collection, vdW tables from global memory.
struct PairVariant {
    int gp;
    bool fs, direct, pp, gt;
    bool comb = false;  // k_pair_comb
    bool rel = false;   // k_pair_comb<REL>
    int nt = kPairThreads;  // workgroup size
    bool operator==(const PairVariant &o) const
    {
        return gp == o.gp && fs == o.fs && direct == o.direct && pp == o.pp && gt == o.gt && comb == o.comb &&
               rel == o.rel && nt == o.nt;
    }
};

template <typename R>
static int pair_setup(Ctx *c, PairArgs<R> &a, size_t &lds, PairVariant &v, int lanes)
{
    using R2 = typename R2T<R>::type;
    PairParams &pp = a.pp;
    pp.ntypes = c->prm.ntypes;
    pp.nkeys = c->nrows;  // distinct table rows (the key table maps type pairs to them)
    pp.mg1 = c->max_grid + 1;
    pp.ns1 = c->nsamples + 1;
    for (int d = 0; d < 3; d++) {
        pp.L[d] = c->L[d];
        pp.invL[d] = c->invL[d];
    }
    pp.electro = c->prm.electrostatics ? 1 : 0;
    pp.rc = c->prm.cutoff;
    pp.rvdw = c->prm.vdw_cutoff;
    const double rmax = pp.electro ? std::max(pp.rc, pp.rvdw) : pp.rvdw;
    pp.rmax2 = rmax * rmax;
    pp.rdr_v = 1.0 / (c->prm.vdw_cutoff / (double)(c->max_grid - 4));
    pp.rdr_e = pp.electro ? 1.0 / (c->prm.cutoff / (double)(c->nsamples - 4)) : 0.0;
    pp.scaling = kR4pie0 / c->prm.epsilon;
    pp.r_rvdw = 1.0 / c->prm.vdw_cutoff;
    pp.cap_lane = c->cap_lane;
    a.pos = c->pos;
    a.typ = c->typ;
    a.n = c->natms;
    a.nbr = c->nbr;
    a.cnt = c->cnt;
    a.vdw_g = (const R2 *)c->d_vdw;
    a.erfc_g = (const R2 *)c->d_erfc;
    a.keytab_g = c->d_keytab;
    a.direct_g = (const R *)c->d_vdw_direct;
    a.frc = c->frc;
    a.red = c->red;
    a.gid = c->gid;
    a.ppo = c->pp;
    a.comb_g = (const float4 *)c->d_comb;
    a.comb_row_g = c->d_comb_row;
    a.prel = c->prel;
    for (int d = 0; d < 3; d++) {
        const double h = c->L[d] / (2.0 * c->nc[d]);
        a.hs[d] = (float)h;
        a.hs_lo[d] = (float)(h - (double)a.hs[d]);
    }
    const bool direct = c->vdw_direct;
    if (sizeof(R) == 4 && comb_ok(c)) {
        const int nt = pp.ntypes;
        const bool rel = c->list_coded && c->rel_pairs;
        pp.nkeys = nt * (nt + 1) / 2;  // combined rows
        lds = (rel ? sizeof(float4) * 256 : 0) + sizeof(float4) * (size_t)pp.nkeys * pp.mg1 +
              sizeof(int32_t) * ((nt * nt + 1) & ~1) + sizeof(double) * 10 * (kCombThreads / 64);
        v = PairVariant{lanes, false, false, false, false, true, rel};
        v.nt = comb_threads(c);
        return 0;
    }
    const size_t lds_rest = sizeof(R2) * (pp.electro ? pp.ns1 : 0) + sizeof(int32_t) * ((pp.ntypes * pp.ntypes + 1) & ~1) +
                            (direct ? sizeof(R) * kDirectStride * pp.nkeys : 0) + sizeof(double) * 10 * (kPairThreads / 64);
    // the vdW tables go to LDS when they fit in its 160 KiB with the rest,
    // else they are read from global memory
    const size_t lds_tab = direct ? 0 : sizeof(R2) * (size_t)pp.nkeys * pp.mg1;
    const bool gtab = !direct && lds_tab + lds_rest > 160 * 1024;
    lds = (gtab ? 0 : lds_tab) + lds_rest;
    if (lds > 160 * 1024)
        return fail(c, DLPX_ERR_UNSUPPORTED, "the erfc and key tables exceed the 160 KiB LDS of a CU");
    // per-particle output, direct mode and global tables run with 8 lanes per atom
    v = PairVariant{(gtab || c->pp_now || direct) ? 8 : lanes, c->vdw_force_shift, direct, c->pp_now, gtab};
    return 0;
}

// Calls CALL(GP, FS, DIRECT, PP, GT) with the compile-time values of v.
#define DLPX_PAIR_VARIANTS(v, CALL)                                                          \
    do {                                                                                     \
        if ((v).gt) {                                                                        \
            if ((v).pp) { if ((v).fs) CALL(8, true, false, true, true); else CALL(8, false, false, true, true); } \
            else { if ((v).fs) CALL(8, true, false, false, true); else CALL(8, false, false, false, true); }      \
        } else if ((v).pp) {                                                                 \
            if ((v).direct) { if ((v).fs) CALL(8, true, true, true, false); else CALL(8, false, true, true, false); } \
            else { if ((v).fs) CALL(8, true, false, true, false); else CALL(8, false, false, true, false); }         \
        } else if ((v).direct) {                                                             \
            if ((v).fs) CALL(8, true, true, false, false); else CALL(8, false, true, false, false); \
        } else if ((v).fs) {                                                                 \
            switch ((v).gp) {                                                                \
                case 16: CALL(16, true, false, false, false); break;                         \
                case 8: CALL(8, true, false, false, false); break;                           \
                default: CALL(4, true, false, false, false); break;                          \
            }                                                                                \
        } else {                                                                             \
            switch ((v).gp) {                                                                \
                case 16: CALL(16, false, false, false, false); break;                        \
                case 8: CALL(8, false, false, false, false); break;                          \
                default: CALL(4, false, false, false, false); break;                         \
            }                                                                                \
        }                                                                                    \
    } while (0)

template <typename R>
static int launch_pair(Ctx *c)
{
    PairArgs<R> a;
    size_t lds;
    PairVariant v;
    int r;
    if ((r = pair_setup<R>(c, a, lds, v, c->pair_lanes))) return r;
    if constexpr (sizeof(R) == 4) {
        if (v.comb) {
            if (v.rel)
                hipLaunchKernelGGL(k_prel, dim3(ceil_div(c->natms, 256)), dim3(256), 0, c->stream, prel_args(c));
#define DLPX_PAIRC(GP, REL, NT)                                                                               \
    do {                                                                                                      \
        DLPX_LDS_MAX(c, (k_pair_comb<GP, NT, REL>));                                                          \
        hipLaunchKernelGGL((k_pair_comb<GP, NT, REL>), dim3(pair_blocks(c)), dim3(NT), lds, c->stream, a);    \
    } while (0)
            if (v.nt == 512) {
                if (v.rel) DLPX_PAIRC(8, true, 512);
                else DLPX_PAIRC(8, false, 512);
            } else {
                switch (v.gp * 2 + (v.rel ? 1 : 0)) {
                    case 33: DLPX_PAIRC(16, true, kCombThreads); break;
                    case 32: DLPX_PAIRC(16, false, kCombThreads); break;
                    case 17: DLPX_PAIRC(8, true, kCombThreads); break;
                    case 16: DLPX_PAIRC(8, false, kCombThreads); break;
                    case 9: DLPX_PAIRC(4, true, kCombThreads); break;
                    default: DLPX_PAIRC(4, false, kCombThreads); break;
                }
            }
#undef DLPX_PAIRC
            DLPX_HIP(c, hipGetLastError());
            return 0;
        }
    }
#define DLPX_PAIR1(GP, FS, DIR, PPK, GTK)                                                                \
    do {                                                                                                 \
        DLPX_LDS_MAX(c, (k_pair<R, GP, kPairThreads, FS, DIR, PPK, GTK>));                                \
        hipLaunchKernelGGL((k_pair<R, GP, kPairThreads, FS, DIR, PPK, GTK>), dim3(pair_blocks(c)),         \
                           dim3(kPairThreads), lds, c->stream, a);                                       \
    } while (0)
    DLPX_PAIR_VARIANTS(v, DLPX_PAIR1);
#undef DLPX_PAIR1
    DLPX_HIP(c, hipGetLastError());
    return 0;
}

int pair_forces(Ctx *c)
{
    if (c->prm.precision == DLPX_MIXED) return launch_pair<float>(c);
    return launch_pair<double>(c);
}

// Batched pair forces: one launch per variant present among the systems
// (every system of the benchmark ensembles shares one).
template <typename R>
static int pair_prepare_t(const std::vector<Ctx *> &cs, PairBatch &pb, int lanes, std::vector<int> &rows)
{
    std::vector<PairVariant> vs;
    std::vector<std::vector<PairArgs<R>>> args;
    std::vector<std::vector<int>> nbs;
    std::vector<size_t> lds_max;
    int r;
    for (Ctx *c : cs) {
        PairArgs<R> a;
        size_t lds;
        PairVariant v;
        if ((r = pair_setup<R>(c, a, lds, v, lanes))) return r;
        size_t g = 0;
        while (g < vs.size() && !(vs[g] == v)) g++;
        if (g == vs.size()) {
            vs.push_back(v);
            args.emplace_back();
            nbs.emplace_back();
            lds_max.push_back(0);
        }
        args[g].push_back(a);
        nbs[g].push_back(ceil_div((long)c->natms * v.gp, v.nt));
        rows.push_back(nbs[g].back());
        lds_max[g] = std::max(lds_max[g], lds);
    }
    pb.groups.resize(vs.size());
    for (size_t g = 0; g < vs.size(); g++) {
        PairBatch::Group &gr = pb.groups[g];
        gr.gp = vs[g].gp;
        gr.fs = vs[g].fs;
        gr.direct = vs[g].direct;
        gr.pp = vs[g].pp;
        gr.gt = vs[g].gt;
        gr.comb = vs[g].comb;
        gr.rel = vs[g].rel;
        gr.nt = vs[g].nt;
        if ((r = batch_upload(cs[0], gr.bl, args[g].data(), sizeof(PairArgs<R>), nbs[g], true))) return r;
        gr.bl.lds = lds_max[g];
    }
    // relative positions for the systems of REL groups
    std::vector<PrelArgs> pa;
    std::vector<int> pnb;
    for (size_t k = 0; k < cs.size(); k++) {
        Ctx *c = cs[k];
        PairArgs<R> a;
        size_t lds;
        PairVariant v;
        if ((r = pair_setup<R>(c, a, lds, v, lanes))) return r;
        pa.push_back(prel_args(c));
        pnb.push_back(v.rel ? ceil_div(c->natms, 256) : 0);
    }
    return batch_upload(cs[0], pb.prel, pa.data(), sizeof(PrelArgs), pnb, false);
}

int pair_prepare(const std::vector<Ctx *> &cs, PairBatch &pb, int lanes, std::vector<int> &rows)
{
    rows.clear();
    if (cs[0]->prm.precision == DLPX_MIXED) return pair_prepare_t<float>(cs, pb, lanes, rows);
    return pair_prepare_t<double>(cs, pb, lanes, rows);
}

template <typename R>
static int pair_launch_t(Ctx *c, hipStream_t s, const PairBatch &pb)
{
    if (pb.prel.total > 0)
        hipLaunchKernelGGL(k_prel_batch, dim3(pb.prel.total), dim3(256), 0, s, (const PrelArgs *)pb.prel.args,
                           pb.prel.boff, pb.prel.nsys);
    for (const PairBatch::Group &gr : pb.groups) {
        PairVariant v{gr.gp, gr.fs, gr.direct, gr.pp, gr.gt, gr.comb, gr.rel};
        v.nt = gr.nt;
        const size_t lds = gr.bl.lds;
        if constexpr (sizeof(R) == 4) {
            if (v.comb) {
#define DLPX_PAIRCB(GP, REL, NT)                                                                                  \
    do {                                                                                                          \
        DLPX_LDS_MAX(c, (k_pair_comb_batch<GP, NT, REL>));                                                         \
        hipLaunchKernelGGL((k_pair_comb_batch<GP, NT, REL>), dim3(gr.bl.total), dim3(NT), lds, s,                  \
                           (const PairArgs<float> *)gr.bl.args, gr.bl.boff, gr.bl.nsys);                          \
    } while (0)
                if (v.nt == 512) {
                    if (v.rel) DLPX_PAIRCB(8, true, 512);
                    else DLPX_PAIRCB(8, false, 512);
                } else {
                    switch (v.gp * 2 + (v.rel ? 1 : 0)) {
                        case 33: DLPX_PAIRCB(16, true, kCombThreads); break;
                        case 32: DLPX_PAIRCB(16, false, kCombThreads); break;
                        case 17: DLPX_PAIRCB(8, true, kCombThreads); break;
                        case 16: DLPX_PAIRCB(8, false, kCombThreads); break;
                        case 9: DLPX_PAIRCB(4, true, kCombThreads); break;
                        default: DLPX_PAIRCB(4, false, kCombThreads); break;
                    }
                }
#undef DLPX_PAIRCB
                continue;
            }
        }
#define DLPX_PAIRB(GP, FS, DIR, PPK, GTK)                                                                   \
    do {                                                                                                    \
        DLPX_LDS_MAX(c, (k_pair_batch<R, GP, kPairThreads, FS, DIR, PPK, GTK>));                             \
        hipLaunchKernelGGL((k_pair_batch<R, GP, kPairThreads, FS, DIR, PPK, GTK>), dim3(gr.bl.total),         \
                           dim3(kPairThreads), lds, s, (const PairArgs<R> *)gr.bl.args, gr.bl.boff,           \
                           gr.bl.nsys);                                                                     \
    } while (0)
        DLPX_PAIR_VARIANTS(v, DLPX_PAIRB);
#undef DLPX_PAIRB
    }
    DLPX_HIP(c, hipGetLastError());
    return 0;
}

int pair_launch(Ctx *c, hipStream_t s, const PairBatch &pb)
{
    if (c->prm.precision == DLPX_MIXED) return pair_launch_t<float>(c, s, pb);
    return pair_launch_t<double>(c, s, pb);
}

int pair_blocks_exact(const Ctx *c)
{
    return ceil_div((long)c->natms * c->pair_lanes, comb_ok(c) ? comb_threads(c) : kPairThreads);
}
int pair_blocks(const Ctx *c) { return xcd_grid(pair_blocks_exact(c)); }

// Ewald correction of the excluded pairs (ewald_spole.F90:479-679, called
// from two_body.F90:551-592).  The reciprocal sum holds the whole Coulomb
// interaction of every pair, so each excluded pair closer than the cutoff
// gives back its smooth part, U = -q_i q_j erf(alpha r)/r, with erf from
// Abramowitz & Stegun 7.1.26 and, below alpha r = 0.01, from its Taylor
// series to x^9.  FP64 in both precision modes: an atom has a handful of
// exclusions.  Thread i (sorted order) walks the exclusion row of atom i and
// adds the forces on i alone, so every pair is met from both ends and its
// energy and stress are halved.
struct ExclParams {
    double L[3], invL[3];
    double rc, alpha, scaling;
};

struct ExclArgs {
    const double4 *pos;
    int n;
    const int32_t *gid, *slot, *excl_off, *excl_ids;
    ExclParams ep;
    double4 *frc;
    double *red;
};
__device__ __forceinline__ void excl_body(const ExclArgs &a, int bid)
{
    // A&S 7.1.26: erfc(x) ~ t (a1 + t (a2 + t (a3 + t (a4 + t a5)))) exp(-x^2), t = 1/(1 + p x)
    constexpr double a1 = 0.254829592, a2 = -0.284496736, a3 = 1.421413741, a4 = -1.453152027,
                     a5 = 1.061405429, p = 0.3275911;
    constexpr double sqrpi = 1.7724538509055160273;
    const double4 *__restrict__ pos = a.pos;
    const int n = a.n;
    const int32_t *__restrict__ gid = a.gid;
    const int32_t *__restrict__ slot = a.slot;
    const int32_t *__restrict__ excl_off = a.excl_off;
    const int32_t *__restrict__ excl_ids = a.excl_ids;
    const ExclParams &ep = a.ep;
    double4 *__restrict__ frc = a.frc;
    __shared__ double lred[7 * (256 / 64)];
    const int i = bid * blockDim.x + threadIdx.x;
    double acc[7] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    if (i < n) {
        const int h = gid[i];
        const int e0 = excl_off[h], e1 = excl_off[h + 1];
        const double4 pi = pos[i];
        if (e1 > e0 && fabs(pi.w) > kZeroPlus) {
            const double chgea = pi.w * ep.scaling, alpha = ep.alpha;
            double fx = 0.0, fy = 0.0, fz = 0.0;
            for (int e = e0; e < e1; e++) {
                const double4 pj = pos[slot[excl_ids[e]]];
                const double dx = min_image_1(pi.x, pj.x, ep.L[0], ep.invL[0]);
                const double dy = min_image_1(pi.y, pj.y, ep.L[1], ep.invL[1]);
                const double dz = min_image_1(pi.z, pj.z, ep.L[2], ep.invL[2]);
                const double rrr = sqrt(dx * dx + dy * dy + dz * dz);
                if (!(fabs(pj.w) > kZeroPlus && rrr < ep.rc)) continue;
                const double chgprd = pj.w * chgea;
                const double rsq = rrr * rrr;
                const double alpr = rrr * alpha, alpr2 = alpr * alpr;
                double erfr, egamma;
                if (alpr < 1.0e-2) {
                    // erf(x)/x = 2/sqrt(pi) (1 - x^2/3 + x^4/10 - x^6/42 + x^8/216), and its r-derivative
                    erfr = 2.0 * chgprd * (alpha / sqrpi) *
                           (1.0 + alpr2 * (-(1.0 / 3.0) + alpr2 * (0.1 + alpr2 * (-(1.0 / 42.0) + alpr2 * (1.0 / 216.0)))));
                    egamma = -4.0 * chgprd * (alpha * alpha * alpha / sqrpi) *
                             ((1.0 / 3.0) + alpr2 * (-2.0 * 0.1 + alpr2 * (3.0 * (1.0 / 42.0) - 4.0 * alpr2 * (1.0 / 216.0))));
                } else {
                    const double ex = exp(-(alpr * alpr));
                    const double tt = 1.0 / (1.0 + p * alpha * rrr);
                    erfr = chgprd * (1.0 - tt * (a1 + tt * (a2 + tt * (a3 + tt * (a4 + tt * a5)))) * ex) / rrr;
                    egamma = -(erfr - 2.0 * chgprd * (alpha / sqrpi) * ex) / rsq;
                }
                const double fxp = egamma * dx, fyp = egamma * dy, fzp = egamma * dz;
                fx += fxp;
                fy += fyp;
                fz += fzp;
                acc[0] -= erfr;
                acc[1] += dx * fxp;
                acc[2] += dx * fyp;
                acc[3] += dx * fzp;
                acc[4] += dy * fyp;
                acc[5] += dy * fzp;
                acc[6] += dz * fzp;
            }
            double4 f = frc[i];
            f.x += fx;
            f.y += fy;
            f.z += fz;
            frc[i] = f;
        }
    }
#pragma unroll
    for (int k = 0; k < 7; k++) acc[k] *= 0.5;  // every pair is met twice
    block_sum<7>(acc, lred, a.red + (size_t)bid * kReduceSlots + R_ENGEX);
}
__global__ void __launch_bounds__(256) k_excl(ExclArgs a) { excl_body(a, blockIdx.x); }
__global__ void __launch_bounds__(256) k_excl_batch(const ExclArgs *__restrict__ av, const int32_t *__restrict__ boff,
                                                    int nsys)
{
    const int lb = blockIdx.x;
    if (lb >= boff[nsys]) return;
    const int s = batch_find(boff, nsys, lb);
    const ExclArgs &a = av[s];
    excl_body(a, lb - boff[s]);
}

int excl_blocks(const Ctx *c) { return ceil_div(c->natms, 256); }

static ExclArgs excl_args(const Ctx *c)
{
    ExclArgs a;
    for (int d = 0; d < 3; d++) {
        a.ep.L[d] = c->L[d];
        a.ep.invL[d] = c->invL[d];
    }
    a.ep.rc = c->prm.cutoff;
    a.ep.alpha = c->prm.alpha;
    a.ep.scaling = kR4pie0 / c->prm.epsilon;
    a.pos = c->pos;
    a.n = c->natms;
    a.gid = c->gid;
    a.slot = c->slot;
    a.excl_off = c->excl_off;
    a.excl_ids = c->excl_ids;
    a.frc = c->frc;
    a.red = c->red;
    return a;
}

int excl_forces(Ctx *c)
{
    if (c->excl_total == 0 || !c->prm.electrostatics) return 0;
    if (c->excl_natms != c->natms) return fail(c, DLPX_ERR_STATE, "exclusion lists are for another atom count");
    hipLaunchKernelGGL(k_excl, dim3(excl_blocks(c)), dim3(256), 0, c->stream, excl_args(c));
    DLPX_HIP(c, hipGetLastError());
    return 0;
}

// Batched exclusion corrections over the systems that have exclusions.
int excl_prepare(const std::vector<Ctx *> &cs, BatchLaunch &bl)
{
    std::vector<ExclArgs> a;
    std::vector<int> nb;
    for (Ctx *c : cs) {
        const bool on = c->excl_total > 0 && c->prm.electrostatics;
        if (on && c->excl_natms != c->natms) return fail(c, DLPX_ERR_STATE, "exclusion lists are for another atom count");
        a.push_back(excl_args(c));
        nb.push_back(on ? excl_blocks(c) : 0);
    }
    return batch_upload(cs[0], bl, a.data(), sizeof(ExclArgs), nb, false);
}

int excl_launch(Ctx *c, hipStream_t s, const BatchLaunch &bl)
{
    if (bl.total == 0) return 0;
    hipLaunchKernelGGL(k_excl_batch, dim3(bl.total), dim3(256), 0, s, (const ExclArgs *)bl.args, bl.boff, bl.nsys);
    DLPX_HIP(c, hipGetLastError());
    return 0;
}

}  // namespace dlpx
