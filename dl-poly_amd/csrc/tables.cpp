// Host-side set-up arithmetic of the force path: interpolation tables,
// B-spline moduli, cell algebra and vdW long-range corrections.  These run
// once per context (or per cell change) on the host, in FP64.
#include <cmath>
#include <vector>

#include "common.h"

namespace dlpx {

static int nint_pos(double x) { return (int)std::floor(x + 0.5); }

// bounds.F90:811,820 -- max(1004, nint(cut/delr_max) + 4), delr_max = 0.01 A
int table_size(double cut) { return std::max(1004, nint_pos(cut / 0.01) + 4); }

// U(r) and gamma(r) = -r dU/dr of the vdW forms the library generates
// (vdw.F90:64-113 keys; two_body_potentials.F90 closed forms):
//   1 12-6       A/r^12 - B/r^6                          (:307-317)
//   2 LJ         4 eps [(sig/r)^12 - (sig/r)^6]          (:260-270)
//   3 n-m        e0/(n-m) [m (r0/r)^n - n (r0/r)^m]      (:330-345; integer powers)
//   4 Buckingham A exp(-r/rho) - C/r^6                   (:471-485)
//   5 BHM        A exp(B (sig-r)) - C/r^6 - D/r^8        (:499-514)
//   6 12-10      A/r^12 - B/r^10                         (:530-546)
//  11 AMOEBA 14-7 eps [1.07/(r/sig + 0.07)]^7 [1.12/((r/sig)^7 + 0.12) - 2]
//                                                        (:657-676)
//   7 n-m shift  n-m shifted and force-corrected at r_trunc (:360-400)
//   8 Morse      e0 {[1 - exp(-k (r-r0))]^2 - 1}         (:416-428)
//   9 WCA        4 eps [(sig/(r-d))^12 - (sig/(r-d))^6] + eps below
//                cut = 2^(1/6) sig + d, else 0           (:559-576)
//  10 DPD        a rc (1 - r/rc)^2 / 2 below rc          (:591-610)
//  12 LJ cohesive 4 eps [(sig/r)^12 - c (sig/r)^6]       (:283-293)
//  13 Morse + 12 e0 {[1 - exp(-k (r-r0))]^2 - 1} + c/r^12 (:442-456)
//  14 Rydberg    (a + b r) exp(-r/c)                     (:689-702)
//  15 ZBL        Z1 Z2 r4pie0/r sum_i b_i exp(-c_i r/a), 1/a = (Z1^0.23 +
//                Z2^0.23)/(0.88534 a_B)                  (:716-741)
//  16 ZBL-Morse  f z + (1-f) Morse, f the Fermi switch at rm of width ic
//                (:754-774 switch, :787-808)
//  17 ZBL-Buck   f z + (1-f) Buckingham                  (:826-847)
//  18 LJ-MDF     LJ(eps, sig) m(r), m the mollified damping from ri to the
//                vdW cutoff (:893-911 damping, :959-976)
//  19 Buck-MDF   Buckingham(A, rho, C) m(r)              (:991-1009)
//  20 12-6-MDF   (a/r^12 - b/r^6) m(r)                   (:1025-1042)
//  21 LJF        ea (rc^2/r^2 - 1)^2 (sig^2/r^2 - 1) below rc (:924-945)
//  22 Sanderson  -A exp(-[(r-L)/d]^2)                    (:865-879)
//  23 nDPD       a rc (1-r/rc)^2 [b (1-r/rc)^(n-1)/(n+1) - 1/2]
//                below rc                                (:623-642)
//  24 Stillinger-Weber two-body part A eps [B (sig/r)^p - (sig/r)^q]
//                exp(sig/(r - aa sig)) below aa sig      (:1058-1082)
// (no long-range correction for 7, 9, 10, 21-24: vdw.F90:681-933)
bool vdw_kind_supported(int kind)
{
    return kind >= 1 && kind <= 24;
}

// The constants the evaluators take, from one key's parameters (the FIELD
// reader's internal units): the ZBL forms carry Z1 Z2 r4pie0 and the inverse
// screening length in slots 0 and 1 (the two quantities zbl_energy forms from
// Z1 and Z2 on every call, two_body_potentials.F90:725-727); the MDF forms
// get the vdW cutoff as the damping's outer radius, as potential_cutoffs
// sets it (vdw.F90:205-230).
void vdw_prepare(int kind, const double *p, double rvdw, double *q)
{
    for (int k = 0; k < kVdwParams; k++) q[k] = p[k];
    if (kind == 15 || kind == 16 || kind == 17) {
        q[0] = p[0] * p[1] * kR4pie0;
        q[1] = (std::pow(p[0], 0.23) + std::pow(p[1], 0.23)) / (0.52917721067 * 0.88534);
    } else if (kind == 18 || kind == 20) {
        q[3] = rvdw;
    } else if (kind == 19) {
        q[4] = rvdw;
    }
}

static double ipow(double x, int n)
{
    double r = 1.0;
    for (int k = 0; k < n; k++) r *= x;
    return r;
}

// ZBL screened Coulomb with kk = Z1 Z2 r4pie0 and a the inverse screening length
static void zbl(double kk, double a, double r, double *e, double *g)
{
    static const double b[4] = {0.18175, 0.50986, 0.28022, 0.02817}, c[4] = {3.1998, 0.94229, 0.40290, 0.20162};
    const double x = r * a;
    double s = 0.0, sc = 0.0;
    for (int i = 0; i < 4; i++) {
        const double t = b[i] * std::exp(-x * c[i]);
        s += t;
        sc += c[i] * t;
    }
    *e = kk * s * (1.0 / r);
    *g = *e + a * kk * sc;
}

// Fermi switch of the ZBL mixtures: 1 - exp(-(rm-r)/ic)/2 below rm,
// exp(-(r-rm)/ic)/2 above
static void fermi(double rm, double ic, double r, double *f, double *df)
{
    const double c = 1.0 / ic;
    const double t = (r < rm ? std::exp(-(rm - r) * c) : std::exp(-(r - rm) * c)) * 0.5;
    *f = r < rm ? 1.0 - t : t;
    *df = r * c * t;
}

// mollified damping function: 1 below ri, 0 above rc, a quintic between
static void mdf(double ri, double rc, double r, double *m, double *dm)
{
    if (r < ri) {
        *m = 1.0;
        *dm = 0.0;
    } else if (r > rc) {
        *m = *dm = 0.0;
    } else {
        const double w = rc - ri, rci = w * w * w * w * w, u = rc - r;
        *m = u * u * u * (10.0 * ri * ri - 5.0 * rc * ri - 15.0 * r * ri + rc * rc + 3.0 * r * rc + 6.0 * r * r) / rci;
        *dm = 30.0 * r * (r - rc) * (r - rc) * (r - ri) * (r - ri) / rci;
    }
}

static void vdw_eval(int kind, const double *p, double r, double *e, double *g)
{
    switch (kind) {
        case 15:  // kk, 1/a
            zbl(p[0], p[1], r, e, g);
            break;
        case 16:  // kk, 1/a, rm, ic, e0, r0, k
        case 17: {  // kk, 1/a, rm, ic, A, rho, C
            double z, dz, f, df, m, dm;
            zbl(p[0], p[1], r, &z, &dz);
            fermi(p[2], p[3], r, &f, &df);
            vdw_eval(kind == 16 ? 8 : 4, p + 4, r, &m, &dm);
            *e = f * z + (1.0 - f) * m;
            *g = f * dz + df * z + (1.0 - f) * dm - df * m;
            break;
        }
        case 18:  // eps, sig, ri, rc
        case 19:  // A, rho, C, ri, rc
        case 20: {  // a, b, ri, rc
            double u, du, m, dm;
            const int base = kind == 18 ? 2 : kind == 19 ? 4 : 1;
            const int nb = kind == 19 ? 3 : 2;
            vdw_eval(base, p, r, &u, &du);
            mdf(p[nb], p[nb + 1], r, &m, &dm);
            *e = u * m;
            *g = du * m + dm * u;
            break;
        }
        case 21: {  // ea, sig^2, rc^2
            const double r2 = r * r;
            *e = *g = 0.0;
            if (r2 <= p[2]) {
                const double ir = 1.0 / r2, st = p[1] * ir, rct = p[2] * ir, x = p[0] * (rct - 1.0) * (rct - 1.0);
                *e = x * (st - 1.0);
                *g = 4.0 * p[0] * rct * (rct - 1.0) * (st - 1.0) + 2.0 * x * st;
            }
            break;
        }
        case 24: {  // eps, A, B, sig, p, q, aa
            *e = *g = 0.0;
            const double rs = p[6] * p[3];
            if (r < rs) {
                const double ex = p[3] / (r - rs), pr = p[3] / r, c = p[2] * std::pow(pr, p[4]);
                const double prq = std::pow(pr, p[5]), xe = std::exp(ex);
                const double t = p[1] * p[0] * (c - prq) * xe;
                *e = t;
                *g = p[1] * p[0] * (p[4] * c - p[5] * prq) * xe + t * r * ex / (r - rs);
            }
            break;
        }
        case 1: {
            const double r6 = ipow(1.0 / r, 6);
            *e = (p[0] * r6 - p[1]) * r6;
            *g = 6.0 * r6 * (2.0 * p[0] * r6 - p[1]);
            break;
        }
        case 2: {
            double s = p[1] / r, s2 = s * s, s6 = s2 * s2 * s2;
            *e = 4.0 * p[0] * s6 * (s6 - 1.0);
            *g = 24.0 * p[0] * s6 * (2.0 * s6 - 1.0);
            break;
        }
        case 3: {
            const double a = p[3] / r, b = 1.0 / (p[1] - p[2]);
            const double rn = ipow(a, (int)p[1]), rm = ipow(a, (int)p[2]);
            *e = p[0] * (p[2] * rn - p[1] * rm) * b;
            *g = p[0] * p[2] * p[1] * (rn - rm) * b;
            break;
        }
        case 4: {
            const double b = r / p[1], t1 = p[0] * std::exp(-b), t2 = -p[2] / ipow(r, 6);
            *e = t1 + t2;
            *g = t1 * b + 6.0 * t2;
            break;
        }
        case 6: {
            const double ri2 = 1.0 / (r * r), f12 = p[0] * ipow(ri2, 6), f10 = -p[1] * ipow(ri2, 5);
            *e = f12 + f10;
            *g = 12.0 * f12 + 10.0 * f10;
            break;
        }
        case 8: {
            const double t = std::exp(-p[2] * (r - p[1]));
            *e = p[0] * ((1.0 - t) * (1.0 - t) - 1.0);
            *g = -2.0 * r * p[0] * p[2] * (1.0 - t) * t;
            break;
        }
        case 9: {  // params eps, sig, d, cut
            *e = *g = 0.0;
            if (r < p[3] || std::fabs(r - p[2]) < 1.0e-10) {
                const double s = p[1] / (r - p[2]), s6 = ipow(s, 6);
                *e = 4.0 * p[0] * s6 * (s6 - 1.0) + p[0];
                *g = 24.0 * p[0] * s6 * (2.0 * s6 - 1.0) * r / (r - p[2]);
            }
            break;
        }
        case 10: {  // params a, rc
            *e = *g = 0.0;
            if (r < p[1]) {
                const double x = r / p[1], h = 0.5 * p[0] * p[1] * (1.0 - x);
                *e = h * (1.0 - x);
                *g = 2.0 * h * x;
            }
            break;
        }
        case 11: {  // params eps, sig
            const double rho = r / p[1], t1 = 1.0 / (0.07 + rho), t2 = 1.0 / (0.12 + std::pow(rho, 7));
            const double t3 = p[0] * std::pow(1.07 * t1, 7), t = t3 * (1.12 * t2 - 2.0);
            *e = t;
            *g = 7.0 * (t1 * t + 1.12 * t3 * t2 * t2 * std::pow(rho, 6)) * rho;
            break;
        }
        case 12: {  // params eps, sig, c
            const double s6 = ipow(p[1] / r, 6);
            *e = 4.0 * p[0] * s6 * (s6 - p[2]);
            *g = 24.0 * p[0] * s6 * (2.0 * s6 - p[2]);
            break;
        }
        case 13: {  // params e0, r0, k, c
            const double t = std::exp(-p[2] * (r - p[1])), c12 = p[3] / ipow(r, 12);
            *e = p[0] * t * (t - 2.0) + c12;
            *g = -2.0 * r * p[0] * p[2] * (1.0 - t) * t + 12.0 * c12;
            break;
        }
        case 14: {  // params a, b, c
            const double x = r / p[2], t = std::exp(-x);
            *e = (p[0] + p[1] * r) * t;
            *g = x * t * (p[0] - p[1] * p[2] + p[1] * r);
            break;
        }
        case 7: {  // params e0, n, m, r0, r_trunc; n and m rounded for the powers
            *e = *g = 0.0;
            if (r <= p[4]) {
                const double n = p[1], m = p[2], ni = std::round(n), mi = std::round(m), t = n - m, b = 1.0 / t;
                const double c = p[4] / p[3], ci = p[3] / p[4];
                const double beta = c * std::pow((std::pow(c, mi + 1.0) - 1.0) / (std::pow(c, ni + 1.0) - 1.0), b);
                const double alpha = -t / (m * std::pow(beta, ni) * (1.0 + (n * ci - n - 1.0) * std::pow(ci, ni)) -
                                           n * std::pow(beta, mi) * (1.0 + (m * ci - m - 1.0) * std::pow(ci, mi)));
                const double e1 = p[0] * alpha, a = p[3] / r;
                const double bn = std::pow(beta * ci, ni) - std::pow(beta * ci, mi);
                *e = e1 * (m * std::pow(beta, ni) * (std::pow(a, ni) - std::pow(ci, ni)) -
                           n * std::pow(beta, mi) * (std::pow(a, mi) - std::pow(ci, mi)) + n * m * ((r / p[4] - 1.0) * bn)) * b;
                *g = e1 * m * n * (std::pow(beta, ni) * std::pow(a, ni) - std::pow(beta, mi) * std::pow(a, mi) - r / p[4] * bn) * b;
            }
            break;
        }
        case 22: {  // params A, L, d
            const double x = (r - p[1]) / p[2], t = p[0] * std::exp(-x * x);
            *e = -t;
            *g = -2.0 * (r - p[1]) * r * t / (p[2] * p[2]);
            break;
        }
        case 23: {  // params a, b, n, rc
            *e = *g = 0.0;
            if (r < p[3]) {
                const double x = r / p[3], t1 = p[0] * p[3] * (1.0 - x), t0 = p[1] * std::pow(1.0 - x, p[2] - 1.0);
                *e = t1 * (1.0 - x) * (t0 / (p[2] + 1.0) - 0.5);
                *g = t1 * x * (t0 - 1.0);
            }
            break;
        }
        default: {  // 5, Born-Huggins-Meyer
            double ri2 = 1.0 / (r * r);
            double t1 = p[0] * std::exp(p[1] * (p[2] - r));
            double t2 = -p[3] * (ri2 * ri2 * ri2);
            double t3 = -p[4] * (ri2 * ri2 * ri2 * ri2);
            *e = t1 + t2 + t3;
            *g = t1 * r * p[1] + 6.0 * t2 + 8.0 * t3;
        }
    }
}

void vdw_form(int kind, const double *q, double r, double *e, double *g) { vdw_eval(kind, q, r, e, g); }

// vdw.F90:1397-1451: samples at r_i = i * rvdw/(max_grid-4), i = 1..max_grid;
// slot 0 holds Huge (never used inside the cutoff).  p: the key's parameters.
void vdw_table(int kind, const double *p, double rvdw, int max_grid, double *pot, double *frc)
{
    double q[kVdwParams];
    vdw_prepare(kind, p, rvdw, q);
    const double dlrpot = rvdw / (double)(max_grid - 4);
    for (int i = 1; i <= max_grid; i++) vdw_eval(kind, q, (double)i * dlrpot, &pot[i], &frc[i]);
    pot[0] = frc[0] = 1.79769313486231570815e308;
}

// numerics.F90:3647-3667: the Abramowitz & Stegun 7.1.26 form of erfc that the
// reference tabulates (not the exact erfc).
static double erfc_as(double x)
{
    const double a1 = 0.254829592, a2 = -0.284496736, a3 = 1.421413741, a4 = -1.453152027,
                 a5 = 1.061405429, pp = 0.3275911;
    const double tt = 1.0 / (1.0 + pp * x);
    return tt * (a1 + tt * (a2 + tt * (a3 + tt * (a4 + tt * a5)))) * std::exp(-(x * x));
}

// electrostatic.F90:88-127 with numerics.F90:215-248 and 3669-3683:
// g(r) = erfc(alpha r)/r and g'(r) = [g(r) + alpha 2/sqrt(pi) exp(-alpha^2 r^2)]/r^2,
// value i-1 at r = i * rc/(n-4).
void erfc_tables(double rc, double alpha, int n, double *erfc_t, double *erfc_d)
{
    const double rsqrpi = 1.0 / std::sqrt(kPi);
    const double dr = rc / (double)(n - 4);
    for (int i = 1; i <= n; i++) {
        const double r = (double)i * dr;
        const double g = erfc_as(alpha * r) / r;
        erfc_t[i - 1] = g;
        const double ar = alpha * r;
        erfc_d[i - 1] = (g + alpha * (2.0 * std::exp(-(ar * ar)) * rsqrpi)) / (r * r);
    }
}

// bspline.F90:73-190: |b(m)|^2 of the Euler exponential spline,
// b(m) = exp(2 pi i (S-1) m/K) / sum_{k=0}^{S-2} M_S(k+1) exp(2 pi i m k/K).
void bspline_norm2(int K, int S, double *norm2)
{
    std::vector<double> wr(K), wi(K);
    wr[0] = 1.0;
    wi[0] = 0.0;
    for (int i = 1; i <= K / 2; i++) {
        const double arg = (2.0 * kPi / (double)K) * (double)i;
        wr[i] = std::cos(arg);
        wi[i] = std::sin(arg);
        wr[K - i] = wr[i];
        wi[K - i] = -wi[i];
    }
    // cardinal B-spline of order S at the integer knots
    std::vector<double> m(S + 2, 0.0);
    m[2] = 1.0;
    for (int k = 3; k <= S; k++)
        for (int j = k; j >= 2; j--)
            m[j] = ((double)(j - 1) * m[j] + (double)(k - j + 1) * m[j - 1]) * (1.0 / (double)(k - 1));
    for (int i = 0; i < K; i++) {
        double tr = 0.0, ti = 0.0;
        for (int k = 0; k <= S - 2; k++) {
            const int e = (int)(((long)i * k) % K);
            tr += m[k + 2] * wr[e];
            ti += m[k + 2] * wi[e];
        }
        const int e = (int)(((long)i * (S - 1)) % K);
        const double den = tr * tr + ti * ti;
        const double br = (wr[e] * tr + wi[e] * ti) / den;
        const double bi = (wi[e] * tr - wr[e] * ti) / den;
        norm2[i] = br * br + bi * bi;
    }
}

// numerics.F90:1450-1509 (inverse by cofactors) and 1344-1448 (cell widths)
void invert_cell(const double *a, double *b, double *det)
{
    b[0] = a[4] * a[8] - a[5] * a[7];
    b[1] = a[2] * a[7] - a[1] * a[8];
    b[2] = a[1] * a[5] - a[2] * a[4];
    b[3] = a[5] * a[6] - a[3] * a[8];
    b[4] = a[0] * a[8] - a[2] * a[6];
    b[5] = a[2] * a[3] - a[0] * a[5];
    b[6] = a[3] * a[7] - a[4] * a[6];
    b[7] = a[1] * a[6] - a[0] * a[7];
    b[8] = a[0] * a[4] - a[1] * a[3];
    const double d = a[0] * b[0] + a[3] * b[1] + a[6] * b[2];
    const double r = std::fabs(d) > 0.0 ? 1.0 / d : 0.0;
    for (int i = 0; i < 9; i++) b[i] *= r;
    *det = d;
}

void dcell(const double *a, double *b)
{
    b[0] = std::sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]);
    b[1] = std::sqrt(a[3] * a[3] + a[4] * a[4] + a[5] * a[5]);
    b[2] = std::sqrt(a[6] * a[6] + a[7] * a[7] + a[8] * a[8]);
    b[3] = (a[0] * a[3] + a[1] * a[4] + a[2] * a[5]) / (b[0] * b[1]);
    b[4] = (a[0] * a[6] + a[1] * a[7] + a[2] * a[8]) / (b[0] * b[2]);
    b[5] = (a[3] * a[6] + a[4] * a[7] + a[5] * a[8]) / (b[1] * b[2]);
    const double axb[3] = {a[1] * a[5] - a[2] * a[4], a[2] * a[3] - a[0] * a[5], a[0] * a[4] - a[1] * a[3]};
    const double bxc[3] = {a[4] * a[8] - a[5] * a[7], a[5] * a[6] - a[3] * a[8], a[3] * a[7] - a[4] * a[6]};
    const double cxa[3] = {a[7] * a[2] - a[8] * a[1], a[8] * a[0] - a[6] * a[2], a[6] * a[1] - a[7] * a[0]};
    b[9] = std::fabs(a[0] * bxc[0] + a[1] * bxc[1] + a[2] * bxc[2]);
    const double d[3] = {b[9] / std::sqrt(bxc[0] * bxc[0] + bxc[1] * bxc[1] + bxc[2] * bxc[2]),
                         b[9] / std::sqrt(cxa[0] * cxa[0] + cxa[1] * cxa[1] + cxa[2] * cxa[2]),
                         b[9] / std::sqrt(axb[0] * axb[0] + axb[1] * axb[1] + axb[2] * axb[2])};
    const double x[3] = {std::fabs(a[0]) / b[0], std::fabs(a[3]) / b[1], std::fabs(a[6]) / b[2]};
    const double y[3] = {std::fabs(a[1]) / b[0], std::fabs(a[4]) / b[1], std::fabs(a[7]) / b[2]};
    // widths ordered to follow the axis each cell vector is most aligned with
    int ix, i1, i2;
    if (x[0] >= x[1] && x[0] >= x[2]) { ix = 0; i1 = 1; i2 = 2; }
    else if (x[1] >= x[0] && x[1] >= x[2]) { ix = 1; i1 = 0; i2 = 2; }
    else { ix = 2; i1 = 0; i2 = 1; }
    b[6] = d[ix];
    if (y[i1] >= y[i2]) { b[7] = d[i1]; b[8] = d[i2]; } else { b[7] = d[i2]; b[8] = d[i1]; }
}

// vdw.F90:617-967 (the closed-form tails of the generated forms; WCA and DPD
// have none)
// AMOEBA 14-7 tail integrals as vdw.F90:522-613 evaluates them: Simpson
// panels of width 2h (h = rw/(r0 n), n = 10000) over chunks of n steps in the
// reduced distance x = r/r0, from rw/r0 onwards, chunk after chunk until one
// changes the sum by at most prec (3 prec/h), each chunk starting again at
// the previous chunk's last panel (that panel counts twice, as there).
static double mm3(double x, double A, double B, double eps)
{
    const double ia = 1.0 / (A + x), ib = 1.0 / (B + std::pow(x, 7));
    return eps * std::pow((1.0 + A) * ia, 7) * ((1.0 + B) * ib - 2.0);
}
static double dmm3(double x, double r0, double A, double B, double eps)
{
    const double ia = 1.0 / (A + x), ib = 1.0 / (B + std::pow(x, 7));
    return -7.0 * (mm3(x, A, B, eps) * ia - std::pow((1.0 + A) * ia, 7) * ((1.0 + B) * ib * ib) * std::pow(x, 6)) / r0;
}
static double mm3_tail(double r0, double A, double B, double eps, double rw, double prec, bool deriv)
{
    const int n = 10000;
    const double xs = rw / r0, h = (2.0 * xs - xs) / (double)n;
    double s = 0.0, sold = HUGE_VAL;
    for (int j = 1; std::fabs(s - sold) * h / 3.0 > prec; j++) {
        sold = s;
        for (int i = (j - 1) * n; i <= j * n; i += 2) {
            const double x = xs + i * h, x1 = x + h, x2 = x + 2.0 * h;
            if (deriv)
                s += x * x * x * dmm3(x, r0, A, B, eps) + 4.0 * x1 * x1 * x1 * dmm3(x1, r0, A, B, eps) +
                     x2 * x2 * x2 * dmm3(x2, r0, A, B, eps);
            else
                s += x * x * mm3(x, A, B, eps) + 4.0 * x1 * x1 * mm3(x1, A, B, eps) + x2 * x2 * mm3(x2, A, B, eps);
        }
    }
    return deriv ? s * h * std::pow(r0, 4) / 3.0 : s * h * std::pow(r0, 3) / 3.0;
}

// The ZBL and MDF tail integrals as vdw.F90:288-520 evaluates them: the same
// chunked Simpson rule in r itself (h = rw/n from rw, weight r^2 for the
// energy and for gamma alike).  A cap on the chunks keeps parameters whose
// integrand never settles from looping for ever.
template <typename F>
static double tail_simpson(F f, double rw, double prec)
{
    const int n = 10000;
    const double h = (2.0 * rw - rw) / (double)n;
    double s = 0.0, sold = HUGE_VAL;
    for (int j = 1; std::fabs(s - sold) * h / 3.0 > prec && j <= 100000; j++) {
        sold = s;
        for (int i = (j - 1) * n; i <= j * n; i += 2) {
            const double x = rw + i * h, x1 = x + h, x2 = x + 2.0 * h;
            s += x * x * f(x) + 4.0 * (x1 * x1) * f(x1) + (x2 * x2) * f(x2);
        }
    }
    return s * h / 3.0;
}

// energy and gamma tails of one prepared form
static void tail_pair(int kind, const double *q, double rw, double *eadd, double *padd)
{
    *eadd = tail_simpson([&](double x) { double e, g; vdw_eval(kind, q, x, &e, &g); return e; }, rw, 1.0e-12);
    *padd = tail_simpson([&](double x) { double e, g; vdw_eval(kind, q, x, &e, &g); return g; }, rw, 1.0e-12);
}

void vdw_lrc(int ntypes, const double *num_type, const int *kinds, const double *params, double rvdw,
             double volm, double *elrc, double *vlrc, const double *num_frz)
{
    double plrc = 0.0, e = 0.0;
    int key = 0;
    for (int i = 1; i <= ntypes; i++)
        for (int j = 1; j <= i; j++, key++) {
            const double *p = params + kVdwParams * (size_t)key, r = rvdw;
            double eadd = 0.0, padd = 0.0;
            switch (kinds[key]) {
                case 1:
                    eadd = p[0] / (9.0 * std::pow(r, 9)) - p[1] / (3.0 * std::pow(r, 3));
                    padd = 12.0 * p[0] / (9.0 * std::pow(r, 9)) - 6.0 * p[1] / (3.0 * std::pow(r, 3));
                    break;
                case 2:
                    eadd = 4.0 * p[0] * (std::pow(p[1], 12) / (9.0 * std::pow(r, 9)) - std::pow(p[1], 6) / (3.0 * std::pow(r, 3)));
                    padd = 8.0 * p[0] * (6.0 * std::pow(p[1], 12) / (9.0 * std::pow(r, 9)) - std::pow(p[1], 6) / std::pow(r, 3));
                    break;
                case 3: {
                    const double e0 = p[0], nr = p[1], mr = p[2], r0 = p[3];
                    eadd = e0 / (nr - mr) *
                           (mr * std::pow(r0, nr) / ((nr - 3.0) * std::pow(r, nr - 3.0)) -
                            nr * std::pow(r0, mr) / ((mr - 3.0) * std::pow(r, mr - 3.0)));
                    padd = e0 / (nr - mr) * nr * mr *
                           (std::pow(r0, nr) / ((nr - 3.0) * std::pow(r, nr - 3.0)) -
                            std::pow(r0, mr) / ((mr - 3.0) * std::pow(r, mr - 3.0)));
                    break;
                }
                case 4:
                    eadd = -p[2] / (3.0 * std::pow(r, 3));
                    padd = -2.0 * p[2] / std::pow(r, 3);
                    break;
                case 5:
                    eadd = -p[3] / (3.0 * std::pow(r, 3)) - p[4] / (5.0 * std::pow(r, 5));
                    padd = -2.0 * p[3] / std::pow(r, 3) - 8.0 * p[4] / (5.0 * std::pow(r, 5));
                    break;
                case 6:
                    eadd = p[0] / (9.0 * std::pow(r, 9)) - p[1] / (7.0 * std::pow(r, 7));
                    padd = 12.0 * p[0] / (9.0 * std::pow(r, 9)) - 10.0 * p[1] / (7.0 * std::pow(r, 7));
                    break;
                case 11:  // vdw.F90:784-796
                    eadd = mm3_tail(p[1], 0.07, 0.12, p[0], r, 1.0e-12, false);
                    padd = -mm3_tail(p[1], 0.07, 0.12, p[0], r, 1.0e-12, true);
                    break;
                case 12:  // vdw.F90:798-808
                    eadd = 4.0 * p[0] * (std::pow(p[1], 12) / (9.0 * std::pow(r, 9)) -
                                         p[2] * std::pow(p[1], 6) / (3.0 * std::pow(r, 3)));
                    padd = 8.0 * p[0] * (6.0 * std::pow(p[1], 12) / (9.0 * std::pow(r, 9)) -
                                         p[2] * std::pow(p[1], 6) / std::pow(r, 3));
                    break;
                case 14: {  // vdw.F90:832-845: tail integrals of (a + b r) exp(-r/c)
                    const double a = p[0], b = p[1], cc = p[2], t = std::exp(-r / cc);
                    eadd = (b * cc * r * r * r + (3 * b * cc * cc + a * cc) * r * r + (6 * b * cc * cc * cc + 2 * a * cc * cc) * r +
                            6 * b * cc * cc * cc * cc + 2 * a * cc * cc * cc) * t;
                    padd = (b * r * r * r * r + (3 * b * cc + a) * r * r * r + (9 * b * cc * cc + 3 * a * cc) * r * r +
                            (18 * b * cc * cc * cc + 6 * a * cc * cc) * r + 18 * b * cc * cc * cc * cc + 6 * a * cc * cc * cc) * t;
                    break;
                }
                case 8:
                case 13:
                case 16: {  // Morse; Morse + c/r^12 adds that term's tail (vdw.F90:810-830); ZBL-Morse
                            // keeps only the Morse tail (:859-876)
                    const double *pm = kinds[key] == 16 ? p + 4 : p;
                    const double e0 = pm[0], r0 = pm[1], kk = pm[2];
                    if (kk > 2.2250738585072014e-308) {  // Tiny(kk)
                        const double t = std::exp(-kk * (r - r0)), k3 = kk * kk * kk, kr = kk * r;
                        eadd = -2.0 * e0 * t / k3 * ((kr + 1) * (kr + 1) + 1) +
                               e0 * t * t / (4.0 * k3) * ((kr + 1) * (kr + 1) + kk * kk * r * r);
                        padd = -2.0 * e0 * t / k3 * (kr * kr * kr + 3 * kr * kr + 6 * kr + 6) +
                               e0 * t * t / (4.0 * k3) * (4.0 * kr * kr * kr + 6 * kr * kr + 6 * kr + 3);
                        if (kinds[key] == 13) {
                            const double s9 = p[3] / (9.0 * std::pow(r, 9));
                            eadd += s9;
                            padd += 12.0 * s9;
                        }
                    }
                    break;
                }
                case 15: {
                    // vdw.F90:846-857: the reference hands Z1 Z2 r4pie0 and 1/a
                    // to intRadZBL, which sets them as the Z1 and Z2 of a ZBL
                    // object (:288-330), so its integrand is the ZBL of those
                    // two numbers; the same integrand here
                    double q1[kVdwParams], q2[kVdwParams];
                    vdw_prepare(15, p, r, q1);
                    vdw_prepare(15, q1, r, q2);
                    tail_pair(15, q2, r, &eadd, &padd);
                    break;
                }
                case 17: {  // vdw.F90:878-890: the Buckingham tail, exponential term included
                    const double A = p[4], r0 = p[5], cc = p[6], t = A * std::exp(-r / r0);
                    eadd = (r * r + 2 * r0 * r + 2 * r0 * r0) * t * r0 - cc / (3.0 * r * r * r);
                    padd = (r * r * r + 3 * r0 * r * r + 6 * r0 * r0 * r + 6 * r0 * r0 * r0) * t - 2.0 * cc / (r * r * r);
                    break;
                }
                case 18:
                case 19:
                case 20: {
                    // vdw.F90:892-921 with intRadMDF (:376-447): the damping's
                    // outer radius is the cutoff the integral starts from, so the
                    // tail is zero unless ri lies beyond it; for LJ-MDF the call
                    // passes eps and sig in each other's places (:403-407)
                    double q[kVdwParams] = {p[0], p[1], p[2], p[3], p[4], p[5], p[6]};
                    if (kinds[key] == 18) {
                        q[0] = p[1];
                        q[1] = p[0];
                    }
                    q[kinds[key] == 19 ? 4 : 3] = r;
                    tail_pair(kinds[key], q, r, &eadd, &padd);
                    break;
                }
                default:  // 7, 9, 10, 21-24: none
                    break;
            }
            if (i != j) {
                eadd *= 2.0;
                padd *= 2.0;
            }
            // frozen-frozen pairs carry no correction (vdw.F90:665-670,953)
            const double frz = num_frz ? num_frz[i - 1] * num_frz[j - 1] : 0.0;
            const double denprd = 2.0 * kPi * (num_type[i - 1] * num_type[j - 1] - frz) / (volm * volm);
            e += volm * denprd * eadd;
            plrc += denprd * padd / 3.0;
        }
    *elrc = e;
    *vlrc = plrc * (-3.0 * volm);
}

}  // namespace dlpx
