/*
 * dlpx_oracle.c -- TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C, scalar, double-precision restatement of DL_POLY 5.1's per-step
 * force path (link-cell pair set, tabulated vdW, erfc-damped real-space
 * Coulomb, SPME reciprocal space, NVE velocity Verlet).  It is the checker
 * the GPU path is compared against; it is never linked into, loaded by or
 * called from the product library.  Each function names the reference lines
 * it restates.
 *
 * Pinning (what fixes each function independently of this file):
 *   orc_vdw_table      closed forms of LJ/BHM; test_vdw.F90:46-59 known answers
 *                      (keys 1-24); the reference's own tables and vdW long-range
 *                      corrections (dumped at full precision) for keys 1-24
 *   orc_erfc_tables    A&S 7.1.26 coefficients; exact Gaussian derivative
 *   orc_interp3        exact for quadratic tables
 *   orc_pairs_brute    O(N^2) minimum-image search (the link-cell set is
 *                      compared with it)
 *   orc_bspline        partition of unity, derivative sums, symmetry, finite
 *                      differences, closed-form cubic B-spline
 *   orc_fft3d          naive DFT on small 2.3.5 sizes
 *   orc_excl_pair      erf(alpha r)/r and its derivative against the C
 *                      library's erf (both branches); REVCON/REVIVE of the
 *                      reference for molecules with exclusions
 *   orc_two_body       REVCON forces and REVIVE energies/virials/stress of the
 *                      compiled reference on generated NaCl/argon systems;
 *                      Madelung constant of rocksalt; direct Ewald sum
 *   orc_nve_step       reference trajectories (REVCON/REVIVE after n steps)
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define ORC_PI 3.14159265358979323846
static const double R4PIE0 = 138935.4835;       /* constants.F90:100 */
static const double ZERO_PLUS = 2.2250738585072014e-308; /* Tiny(1.0_wp) */

/* ------------------------------------------------------------------ */
/* cell helpers: numerics.F90:1450-1509 (invert), 1344-1448 (dcell)    */
/* ------------------------------------------------------------------ */
void orc_invert(const double *a, double *b, double *det)
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
    double d = a[0] * b[0] + a[3] * b[1] + a[6] * b[2];
    double r = 0.0;
    if (fabs(d) > 0.0) r = 1.0 / d;
    for (int i = 0; i < 9; i++) b[i] *= r;
    *det = d;
}

void orc_dcell(const double *a, double *b)
{
    b[0] = sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]);
    b[1] = sqrt(a[3] * a[3] + a[4] * a[4] + a[5] * a[5]);
    b[2] = sqrt(a[6] * a[6] + a[7] * a[7] + a[8] * a[8]);
    b[3] = (a[0] * a[3] + a[1] * a[4] + a[2] * a[5]) / (b[0] * b[1]);
    b[4] = (a[0] * a[6] + a[1] * a[7] + a[2] * a[8]) / (b[0] * b[2]);
    b[5] = (a[3] * a[6] + a[4] * a[7] + a[5] * a[8]) / (b[1] * b[2]);
    double axb1 = a[1] * a[5] - a[2] * a[4], axb2 = a[2] * a[3] - a[0] * a[5], axb3 = a[0] * a[4] - a[1] * a[3];
    double bxc1 = a[4] * a[8] - a[5] * a[7], bxc2 = a[5] * a[6] - a[3] * a[8], bxc3 = a[3] * a[7] - a[4] * a[6];
    double cxa1 = a[7] * a[2] - a[8] * a[1], cxa2 = a[8] * a[0] - a[6] * a[2], cxa3 = a[6] * a[1] - a[7] * a[0];
    b[9] = fabs(a[0] * bxc1 + a[1] * bxc2 + a[2] * bxc3);
    double d[3], x[3], y[3];
    d[0] = b[9] / sqrt(bxc1 * bxc1 + bxc2 * bxc2 + bxc3 * bxc3);
    d[1] = b[9] / sqrt(cxa1 * cxa1 + cxa2 * cxa2 + cxa3 * cxa3);
    d[2] = b[9] / sqrt(axb1 * axb1 + axb2 * axb2 + axb3 * axb3);
    x[0] = fabs(a[0]) / b[0]; y[0] = fabs(a[1]) / b[0];
    x[1] = fabs(a[3]) / b[1]; y[1] = fabs(a[4]) / b[1];
    x[2] = fabs(a[6]) / b[2]; y[2] = fabs(a[7]) / b[2];
    if (x[0] >= x[1] && x[0] >= x[2]) {
        b[6] = d[0];
        if (y[1] >= y[2]) { b[7] = d[1]; b[8] = d[2]; } else { b[7] = d[2]; b[8] = d[1]; }
    } else if (x[1] >= x[0] && x[1] >= x[2]) {
        b[6] = d[1];
        if (y[0] >= y[2]) { b[7] = d[0]; b[8] = d[2]; } else { b[7] = d[2]; b[8] = d[0]; }
    } else {
        b[6] = d[2];
        if (y[0] >= y[1]) { b[7] = d[0]; b[8] = d[1]; } else { b[7] = d[1]; b[8] = d[0]; }
    }
}

/* ------------------------------------------------------------------ */
/* potentials and tables                                               */
/* ------------------------------------------------------------------ */
/* U(r) and gamma = -r dU/dr, two_body_potentials.F90; kind = vdw.F90:64-113 key:
 * 1 12-6 (:307-317), 2 LJ (:260-270), 3 n-m (:330-345, integer powers of r0/r),
 * 4 Buckingham (:471-485), 5 BHM (:499-514), 6 12-10 (:530-546), 8 Morse (:416-428).
 * Returns -1 for other kinds. */
static double orc_ipow(double x, int n)
{
    double r = 1.0;
    while (n-- > 0) r *= x;
    return r;
}

int orc_vdw_eval(int kind, const double *p, double r, double *e, double *g)
{
    if (kind == 1) { /* u = a/r^12 - b/r^6 */
        double r6 = orc_ipow(1.0 / r, 6);
        *e = (p[0] * r6 - p[1]) * r6;
        *g = 6.0 * r6 * (2.0 * p[0] * r6 - p[1]);
        return 0;
    }
    if (kind == 2) {
        double sor6 = pow(p[1] / r, 6);
        *e = 4.0 * p[0] * sor6 * (sor6 - 1.0);
        *g = 24.0 * p[0] * sor6 * (2.0 * sor6 - 1.0);
        return 0;
    }
    if (kind == 3) { /* u = e0/(n-m) [m (r0/r)^n - n (r0/r)^m], params e0, n, m, r0 */
        double a = p[3] / r, b = 1.0 / (p[1] - p[2]);
        double rn = orc_ipow(a, (int)p[1]), rm = orc_ipow(a, (int)p[2]);
        *e = p[0] * (p[2] * rn - p[1] * rm) * b;
        *g = p[0] * p[2] * p[1] * (rn - rm) * b;
        return 0;
    }
    if (kind == 4) { /* u = A exp(-r/rho) - C/r^6, params A, rho, C */
        double b = r / p[1], t1 = p[0] * exp(-b), t2 = -p[2] / pow(r, 6);
        *e = t1 + t2;
        *g = t1 * b + 6.0 * t2;
        return 0;
    }
    if (kind == 5) {
        double ri2 = 1.0 / (r * r);
        double t1 = p[0] * exp(p[1] * (p[2] - r));
        double t2 = -p[3] * (ri2 * ri2 * ri2);
        double t3 = -p[4] * (ri2 * ri2 * ri2 * ri2);
        *e = t1 + t2 + t3;
        *g = t1 * r * p[1] + 6.0 * t2 + 8.0 * t3;
        return 0;
    }
    if (kind == 6) { /* u = a/r^12 - b/r^10 */
        double ri2 = 1.0 / (r * r);
        double f12 = p[0] * pow(ri2, 6), f10 = -p[1] * pow(ri2, 5);
        *e = f12 + f10;
        *g = 12.0 * f12 + 10.0 * f10;
        return 0;
    }
    if (kind == 8) { /* u = e0 {[1 - exp(-k (r - r0))]^2 - 1}, params e0, r0, k */
        double t = exp(-p[2] * (r - p[1]));
        *e = p[0] * ((1.0 - t) * (1.0 - t) - 1.0);
        *g = -2.0 * r * p[0] * p[2] * (1.0 - t) * t;
        return 0;
    }
    if (kind == 9) { /* WCA (two_body_potentials.F90:559-576), params eps, sig, d, cut */
        *e = 0.0; *g = 0.0;
        if (r < p[3] || fabs(r - p[2]) < 1.0e-10) {
            double sr6 = pow(p[1] / (r - p[2]), 6);
            *e = 4.0 * p[0] * sr6 * (sr6 - 1.0) + p[0];
            *g = 24.0 * p[0] * sr6 * (2.0 * sr6 - 1.0) * r / (r - p[2]);
        }
        return 0;
    }
    if (kind == 10) { /* DPD (:591-610), params a, rc */
        *e = 0.0; *g = 0.0;
        if (r < p[1]) {
            double t2 = r / p[1], t1 = 0.5 * p[0] * p[1] * (1.0 - t2);
            *e = t1 * (1.0 - t2);
            *g = 2.0 * t1 * t2;
        }
        return 0;
    }
    if (kind == 11) { /* AMOEBA 14-7 (:657-676), params eps, sig */
        double rho = r / p[1], t1 = 1.0 / (0.07 + rho), t2 = 1.0 / (0.12 + pow(rho, 7));
        double t3 = p[0] * pow(1.07 * t1, 7), t = t3 * (1.12 * t2 - 2.0);
        *e = t;
        *g = 7.0 * (t1 * t + 1.12 * t3 * pow(t2, 2) * pow(rho, 6)) * rho;
        return 0;
    }
    if (kind == 12) { /* LJ cohesive (:283-293), params eps, sig, c */
        double sor6 = pow(p[1] / r, 6);
        *e = 4.0 * p[0] * sor6 * (sor6 - p[2]);
        *g = 24.0 * p[0] * sor6 * (2.0 * sor6 - p[2]);
        return 0;
    }
    if (kind == 13) { /* Morse + c/r^12 (:442-456), params e0, r0, k, c */
        double t1 = exp(-p[2] * (r - p[1])), t2 = p[3] * pow(r, -12);
        *e = p[0] * t1 * (t1 - 2.0) + t2;
        *g = -2.0 * r * p[0] * p[2] * (1.0 - t1) * t1 + 12.0 * t2;
        return 0;
    }
    if (kind == 14) { /* Rydberg (:689-702), params a, b, c */
        double kk = r / p[2], t1 = exp(-kk);
        *e = (p[0] + p[1] * r) * t1;
        *g = kk * t1 * (p[0] - p[1] * p[2] + p[1] * r);
        return 0;
    }
    if (kind == 7) { /* n-m shift (:360-400), params e0, n, m, r0, r_trunc; Nint of n, m as powers */
        *e = 0.0; *g = 0.0;
        if (r <= p[4]) {
            double n = p[1], m = p[2], n_int = round(n), m_int = round(m);
            double t = n - m, b = 1.0 / t, c = p[4] / p[3], c_inv = p[3] / p[4];
            double beta = c * pow((pow(c, m_int + 1.0) - 1.0) / (pow(c, n_int + 1.0) - 1.0), b);
            double alpha = -t / (m * pow(beta, n_int) * (1.0 + (n * c_inv - n - 1.0) * pow(c_inv, n_int)) -
                                 n * pow(beta, m_int) * (1.0 + (m * c_inv - m - 1.0) * pow(c_inv, m_int)));
            double e1 = p[0] * alpha, a = p[3] / r;
            *e = e1 * (m * pow(beta, n_int) * (pow(a, n_int) - pow(c_inv, n_int)) -
                       n * pow(beta, m_int) * (pow(a, m_int) - pow(c_inv, m_int)) +
                       n * m * ((r / p[4] - 1.0) * (pow(beta * c_inv, n_int) - pow(beta * c_inv, m_int)))) * b;
            *g = e1 * m * n * (pow(beta, n_int) * pow(a, n_int) - pow(beta, m_int) * pow(a, m_int) -
                               r / p[4] * (pow(beta * c_inv, n_int) - pow(beta * c_inv, m_int))) * b;
        }
        return 0;
    }
    if (kind == 22) { /* Sanderson (:865-879), params A, L, d */
        double b = pow((r - p[1]) / p[2], 2.0), t = p[0] * exp(-b);
        *e = -t;
        *g = -2.0 * (r - p[1]) * r * t / pow(p[2], 2.0);
        return 0;
    }
    if (kind >= 15 && kind <= 17) {
        /* ZBL (:716-741) from Z1, Z2; the ZBL-Morse (:787-808) and ZBL-Buckingham
         * (:826-847) mixtures through the Fermi switch (:754-774) at rm = p[2],
         * width ic = p[3], with the Morse / Buckingham parameters in p[4..6] */
        static const double zb[4] = {0.18175, 0.50986, 0.28022, 0.02817}, zc[4] = {3.1998, 0.94229, 0.40290, 0.20162};
        double a = (pow(p[0], 0.23) + pow(p[1], 0.23)) / (0.52917721067 * 0.88534);
        double kk = p[0] * p[1] * R4PIE0, x = r * a, ir = 1.0 / r, z = 0.0, dz = 0.0;
        for (int i = 0; i < 4; i++) {
            double t1 = zb[i] * exp(-x * zc[i]);
            z += t1;
            dz -= zc[i] * t1;
        }
        z = kk * z * ir;
        dz = z - a * kk * dz;
        if (kind == 15) { *e = z; *g = dz; return 0; }
        double c = 1.0 / p[3], f, df, t, m, dm;
        if (r < p[2]) { t = exp(-(p[2] - r) * c) * 0.5; f = 1.0 - t; df = r * c * t; }
        else { t = exp(-(r - p[2]) * c) * 0.5; f = t; df = r * c * t; }
        orc_vdw_eval(kind == 16 ? 8 : 4, p + 4, r, &m, &dm);
        *e = f * z + (1.0 - f) * m;
        *g = f * dz + df * z + (1.0 - f) * dm - df * m;
        return 0;
    }
    if (kind >= 18 && kind <= 20) {
        /* LJ (:959-976), Buckingham (:991-1009) or 12-6 (:1025-1042) times the
         * mollified damping (:893-911) from ri to rc; params eps, sig, ri, rc /
         * A, rho, C, ri, rc / a, b, ri, rc (rc: the vdW cutoff, orc_vdw_params) */
        int nb = kind == 19 ? 3 : 2;
        double u, du, m, dm, ri = p[nb], rc = p[nb + 1];
        orc_vdw_eval(kind == 18 ? 2 : kind == 19 ? 4 : 1, p, r, &u, &du);
        if (r < ri) { m = 1.0; dm = 0.0; }
        else if (r > rc) { m = 0.0; dm = 0.0; }
        else {
            double rci = pow(rc - ri, 5);
            m = pow(rc - r, 3) * (10.0 * ri * ri - 5.0 * rc * ri - 15.0 * r * ri + rc * rc + 3.0 * r * rc + 6 * r * r) / rci;
            dm = 30.0 * r * pow(r - rc, 2) * pow(r - ri, 2) / rci;
        }
        *e = u * m;
        *g = du * m + dm * u;
        return 0;
    }
    if (kind == 21) { /* LJF (:924-945), params ea, sig^2, rc^2 */
        double r2 = r * r;
        *e = 0.0; *g = 0.0;
        if (!(r2 > p[2])) {
            double ir = 1.0 / r2, st = p[1] * ir, rct = p[2] * ir, x = p[0] * pow(rct - 1.0, 2);
            *e = x * (st - 1.0);
            *g = 4.0 * p[0] * rct * (rct - 1.0) * (st - 1.0) + 2.0 * x * st;
        }
        return 0;
    }
    if (kind == 24) { /* Stillinger-Weber two-body part (:1058-1082), params eps, A, B, sig, p, q, aa */
        double ee = p[3] / (r - p[6] * p[3]);
        *e = 0.0; *g = 0.0;
        if (r < p[6] * p[3]) {
            double p_r = p[3] / r, c = p[2] * pow(p_r, p[4]), exp_e = exp(ee);
            double t = p[1] * p[0] * (c - pow(p_r, p[5])) * exp_e;
            *e = t;
            *g = p[1] * p[0] * (p[4] * c - p[5] * pow(p_r, p[5])) * exp_e + t * r * ee / (r - p[6] * p[3]);
        }
        return 0;
    }
    if (kind == 23) { /* nDPD (:623-642), params a, b, n, rc */
        *e = 0.0; *g = 0.0;
        if (r < p[3]) {
            double t2 = r / p[3], t1 = p[0] * p[3] * (1.0 - t2), t0 = p[1] * pow(1.0 - t2, p[2] - 1.0);
            *e = t1 * (1.0 - t2) * (t0 / (p[2] + 1.0) - 0.5);
            *g = t1 * t2 * (t0 - 1.0);
        }
        return 0;
    }
    return -1;
}

/* vdw.F90:1397-1576: tables at r_i = i*dlrpot, i = 1..max_grid,
 * dlrpot = rvdw/(max_grid-4); index 0 holds Huge. Arrays have max_grid+1 slots. */
/* vdw.F90:205-230 (potential_cutoffs): the damped forms take the vdW cutoff
 * as the damping's outer radius; q gets the key's 7 parameters with it */
void orc_vdw_params(int kind, const double *p, double rvdw, double *q)
{
    memcpy(q, p, 7 * sizeof(double));
    if (kind == 18 || kind == 20) q[3] = rvdw;
    if (kind == 19) q[4] = rvdw;
}

int orc_vdw_table(int kind, const double *p, double rvdw, int max_grid, double *pot, double *frc)
{
    double dlrpot = rvdw / (double)(max_grid - 4), q[7];
    orc_vdw_params(kind, p, rvdw, q);
    for (int i = 1; i <= max_grid; i++) {
        double r = (double)i * dlrpot, e, g;
        if (orc_vdw_eval(kind, q, r, &e, &g)) return -1;
        pot[i] = e;
        frc[i] = g;
    }
    pot[0] = 1.79769313486231570815e308;
    frc[0] = 1.79769313486231570815e308;
    return 0;
}

/* numerics.F90:3647-3667 (A&S 7.1.26 erfc) and 3669-3683 (2/sqrt(pi) exp(-x^2)) */
double orc_erfc_as(double x)
{
    const double a1 = 0.254829592, a2 = -0.284496736, a3 = 1.421413741,
                 a4 = -1.453152027, a5 = 1.061405429, pp = 0.3275911;
    double tt = 1.0 / (1.0 + pp * x);
    return tt * (a1 + tt * (a2 + tt * (a3 + tt * (a4 + tt * a5)))) * exp(-(x * x));
}

/* electrostatic.F90:88-127 + numerics.F90:215-248: tab[i-1] = f(i*dr), dr = rc/(n-4) */
void orc_erfc_tables(double rc, double alpha, int n, double *erfc_t, double *erfc_d)
{
    const double rsqrpi = 1.0 / sqrt(ORC_PI);
    double dr = rc / (double)(n - 4);
    for (int i = 1; i <= n; i++) {
        double r = (double)i * dr;
        double g = orc_erfc_as(alpha * r) / r;
        erfc_t[i - 1] = g;
        double rsq = r * r;
        erfc_d[i - 1] = (g + alpha * (2.0 * exp(-((alpha * r) * (alpha * r))) * rsqrpi)) / rsq;
    }
}

/* numerics.F90:250-290 three_p_interp on a table where tab(l) = f(l*dr).
 * tab0 points at the value for l = 0 (which may be virtual). */
double orc_interp3(const double *tab0, double rdr, double r)
{
    int l = (int)(r * rdr);
    if (l < 1) return NAN; /* below the first sample: the reference reads outside its table */
    double ppp = r * rdr - (double)l;
    double g0 = tab0[l], g1 = tab0[l + 1], g2 = tab0[l + 2];
    if (l == 0) g0 = g0 * r;
    double t1 = g0 + (g1 - g0) * ppp;
    double t2 = g1 + (g2 - g1) * (ppp - 1.0);
    return t1 + (t2 - t1) * ppp * 0.5;
}

/* vdw.F90:617-967 for the forms of orc_vdw_eval: energy and virial long-range corrections.
 * kinds/params are indexed by pair key (0-based key = hi*(hi-1)/2+lo-1, 1-based types). */
/* vdw.F90:522-613: the AMOEBA tail integrals by the reference's chunked
 * Simpson rule (chunks of n = 10000 steps of h = rw/(r0 n) in x = r/r0 from
 * rw/r0, each restarting at the previous chunk's last panel, until a chunk
 * adds at most 3 prec/h) */
static double orc_mm3(double x, double A, double B, double eps)
{
    double ia = 1.0 / (A + x), ib = 1.0 / (B + pow(x, 7));
    return eps * pow((1.0 + A) * ia, 7) * ((1.0 + B) * ib - 2.0);
}
static double orc_dmm3(double x, double r0, double A, double B, double eps)
{
    double ia = 1.0 / (A + x), ib = 1.0 / (B + pow(x, 7));
    return -7.0 * (orc_mm3(x, A, B, eps) * ia - pow((1.0 + A) * ia, 7) * ((1.0 + B) * pow(ib, 2)) * pow(x, 6)) / r0;
}
static double orc_mm3_int(double r0, double A, double B, double eps, double rw, double prec, int deriv)
{
    int n = 10000, j = 1;
    double is = rw / r0, ie = 2.0 * is, h = (ie - is) / (double)n, s = 0.0, sold = HUGE_VAL;
    while (fabs(s - sold) * h / 3.0 > prec) {
        sold = s;
        for (int i = (j - 1) * n; i <= j * n; i += 2) {
            double x = is + i * h;
            if (deriv)
                s += pow(x, 3) * orc_dmm3(x, r0, A, B, eps) + 4.0 * pow(x + h, 3) * orc_dmm3(x + h, r0, A, B, eps) +
                     pow(x + 2.0 * h, 3) * orc_dmm3(x + 2.0 * h, r0, A, B, eps);
            else
                s += x * x * orc_mm3(x, A, B, eps) + 4.0 * pow(x + h, 2) * orc_mm3(x + h, A, B, eps) +
                     pow(x + 2.0 * h, 2) * orc_mm3(x + 2.0 * h, A, B, eps);
        }
        j++;
    }
    return deriv ? s * h * pow(r0, 4) / 3.0 : s * h * pow(r0, 3) / 3.0;
}

/* vdw.F90:288-520: intRadZBL / intdRadZBL / intRadMDF / intdRadMDF, the same
 * chunked Simpson rule in r (h = rw/n from rw, weight r^2) over U (deriv 0)
 * or gamma (deriv 1) of a form with completed parameters q */
static double orc_rad_int(int kind, const double *q, double rw, double prec, int deriv)
{
    int n = 10000, j = 1;
    double is = rw, ie = 2 * is, h = (ie - is) / (double)n, s = 0.0, sold = HUGE_VAL;
    while (fabs(s - sold) * h / 3.0 > prec && j <= 100000) {
        sold = s;
        for (int i = (j - 1) * n; i <= j * n; i += 2) {
            double x = is + i * h, e0, g0, e1, g1, e2, g2;
            orc_vdw_eval(kind, q, x, &e0, &g0);
            orc_vdw_eval(kind, q, x + h, &e1, &g1);
            orc_vdw_eval(kind, q, x + 2.0 * h, &e2, &g2);
            s += x * x * (deriv ? g0 : e0) + 4.0 * pow(x + h, 2) * (deriv ? g1 : e1) +
                 pow(x + 2.0 * h, 2) * (deriv ? g2 : e2);
        }
        j++;
    }
    return s * h / 3.0;
}

void orc_vdw_lrc(int ntypes, const double *num_type, const int *kinds, const double *params /*7 per key*/,
                 double rvdw, double volm, double *elrc, double *vlrc, const double *num_frz /* NULL: none */)
{
    const double twopi = 2.0 * ORC_PI;
    double plrc = 0.0, e = 0.0;
    int ivdw = 0;
    for (int i = 1; i <= ntypes; i++) {
        for (int j = 1; j <= i; j++) {
            double eadd = 0.0, padd = 0.0, r = rvdw;
            const double *p = params + 7 * ivdw;
            int kd = kinds[ivdw];
            if (kd == 1) { /* vdw.F90:697-707 */
                eadd = p[0] / (9.0 * pow(r, 9)) - p[1] / (3.0 * pow(r, 3));
                padd = 12.0 * p[0] / (9.0 * pow(r, 9)) - 6.0 * p[1] / (3.0 * pow(r, 3));
            } else if (kd == 2) { /* :709-718 */
                double eps = p[0], sig = p[1];
                eadd = 4.0 * eps * (pow(sig, 12) / (9.0 * pow(r, 9)) - pow(sig, 6) / (3.0 * pow(r, 3)));
                padd = 8.0 * eps * (6.0 * pow(sig, 12) / (9.0 * pow(r, 9)) - pow(sig, 6) / pow(r, 3));
            } else if (kd == 3) { /* :720-732 */
                double e0 = p[0], nr = p[1], mr = p[2], r0 = p[3];
                double tn = pow(r0, nr) / ((nr - 3.0) * pow(r, nr - 3.0)), tm = pow(r0, mr) / ((mr - 3.0) * pow(r, mr - 3.0));
                eadd = e0 / (nr - mr) * (mr * tn - nr * tm);
                padd = e0 / (nr - mr) * nr * mr * (tn - tm);
            } else if (kd == 4) { /* :734-742: the dispersion tail only */
                eadd = -p[2] / (3.0 * pow(r, 3));
                padd = -2.0 * p[2] / pow(r, 3);
            } else if (kd == 5) { /* :744-753 */
                double c = p[3], d = p[4];
                eadd = -c / (3.0 * pow(r, 3)) - d / (5.0 * pow(r, 5));
                padd = -2.0 * c / pow(r, 3) - 8.0 * d / (5.0 * pow(r, 5));
            } else if (kd == 6) { /* :755-764 */
                eadd = p[0] / (9.0 * pow(r, 9)) - p[1] / (7.0 * pow(r, 7));
                padd = 12.0 * p[0] / (9.0 * pow(r, 9)) - 10.0 * p[1] / (7.0 * pow(r, 7));
            } else if ((kd == 8 || kd == 13 || kd == 16) && p[kd == 16 ? 6 : 2] > ZERO_PLUS) {
                /* :766-782, 810-830; :859-876 (ZBL-Morse: the Morse tail of p[4..6]) */
                const double *pm = kd == 16 ? p + 4 : p;
                double e0 = pm[0], r0 = pm[1], kk = pm[2], t = exp(-kk * (r - r0)), k3 = kk * kk * kk;
                eadd = -2.0 * e0 * t / k3 * (pow(kk * r + 1, 2) + 1) + e0 * t * t / (4.0 * k3) * (pow(kk * r + 1, 2) + kk * kk * r * r);
                padd = -2.0 * e0 * t / k3 * (pow(kk, 3) * pow(r, 3) + 3 * kk * kk * r * r + 6 * kk * r + 6) +
                       e0 * t * t / (4.0 * k3) * (4.0 * pow(kk, 3) * pow(r, 3) + 6 * kk * kk * r * r + 6 * kk * r + 3);
                if (kd == 13) { double s9 = p[3] / (9.0 * pow(r, 9)); eadd += s9; padd += 12.0 * s9; }
            } else if (kd == 11) { /* :784-796 */
                eadd = orc_mm3_int(p[1], 0.07, 0.12, p[0], r, 1.0e-12, 0);
                padd = -orc_mm3_int(p[1], 0.07, 0.12, p[0], r, 1.0e-12, 1);
            } else if (kd == 12) { /* :798-808 */
                eadd = 4.0 * p[0] * (pow(p[1], 12) / (9.0 * pow(r, 9)) - p[2] * pow(p[1], 6) / (3.0 * pow(r, 3)));
                padd = 8.0 * p[0] * (6.0 * pow(p[1], 12) / (9.0 * pow(r, 9)) - p[2] * pow(p[1], 6) / pow(r, 3));
            } else if (kd == 14) { /* :832-845 */
                double a = p[0], b = p[1], c = p[2], t = exp(-r / c);
                eadd = (b * c * pow(r, 3) + (3 * b * c * c + a * c) * r * r + (6 * b * pow(c, 3) + 2 * a * c * c) * r +
                        6 * b * pow(c, 4) + 2 * a * pow(c, 3)) * t;
                padd = (b * pow(r, 4) + (3 * b * c + a) * pow(r, 3) + (9 * b * c * c + 3 * a * c) * r * r +
                        (18 * b * pow(c, 3) + 6 * a * c * c) * r + 18 * b * pow(c, 4) + 6 * a * pow(c, 3)) * t;
            }
            else if (kd == 15) {
                /* :846-857: kk = Z1 Z2 r4pie0 and the inverse screening length go
                 * to intRadZBL, which stores them as the Z1, Z2 of its ZBL */
                double q[7] = {0}, z1 = p[0], z2 = p[1];
                q[0] = z1 * z2 * R4PIE0;
                q[1] = (pow(z1, 0.23) + pow(z2, 0.23)) / (0.52917721067 * 0.88534);
                eadd = orc_rad_int(15, q, r, 1e-12, 0);
                padd = orc_rad_int(15, q, r, 1e-12, 1);
            } else if (kd == 17) { /* :878-890 */
                double A = p[4], r0 = p[5], c = p[6], t = A * exp(-r / r0);
                eadd = (r * r + 2 * r0 * r + 2 * r0 * r0) * t * r0 - c / (3.0 * pow(r, 3));
                padd = (pow(r, 3) + 3 * r0 * r * r + 6 * r0 * r0 * r + 6 * pow(r0, 3)) * t - 2.0 * c / pow(r, 3);
            } else if (kd >= 18 && kd <= 20) {
                /* :892-921: intRadMDF(pot, a, b, c, ri, rw) with rc = rw; "mlj"
                 * sets LJ sigma = a and epsilon = b (:403-407), a = param(1) */
                double q[7] = {0};
                if (kd == 18) { q[0] = p[1]; q[1] = p[0]; q[2] = p[2]; q[3] = r; }
                if (kd == 19) { q[0] = p[0]; q[1] = p[1]; q[2] = p[2]; q[3] = p[3]; q[4] = r; }
                if (kd == 20) { q[0] = p[0]; q[1] = p[1]; q[2] = p[2]; q[3] = r; }
                eadd = orc_rad_int(kd, q, r, 1e-12, 0);
                padd = orc_rad_int(kd, q, r, 1e-12, 1);
            }
            if (i != j) { eadd *= 2.0; padd *= 2.0; }
            /* vdw.F90:953: frozen-frozen pairs carry no correction */
            double frz = num_frz ? num_frz[i - 1] * num_frz[j - 1] : 0.0;
            double denprd = twopi * (num_type[i - 1] * num_type[j - 1] - frz) / (volm * volm);
            e += volm * denprd * eadd;
            plrc += denprd * padd / 3.0;
            ivdw++;
        }
    }
    *elrc = e;
    *vlrc = plrc * (-3.0 * volm);
}

/* ------------------------------------------------------------------ */
/* pair sets                                                           */
/* ------------------------------------------------------------------ */
/* Minimum-image separation r_ij = x_i - (x_j + s L) for an orthorhombic
 * cell; s is returned in sh (box lengths). */
static inline void min_image(const double *L, const double *xi, const double *xj, double dx[3], int sh[3])
{
    for (int d = 0; d < 3; d++) {
        double s = nearbyint((xi[d] - xj[d]) / L[d]);
        sh[d] = (int)s;
        dx[d] = xi[d] - (xj[d] + s * L[d]);
    }
}

/* Brute force O(N^2): every unordered pair i<j with minimum-image r^2 <= rcut^2.
 * Writes (i, j, s_x, s_y, s_z) where r_ij = x_i - (x_j + s*L).
 * Returns the number of pairs (entries past cap are counted, not written). */
long orc_pairs_brute(int n, const double *cell, const double *xyz, double rcut,
                     int32_t *out, long cap)
{
    double L[3] = {cell[0], cell[4], cell[8]};
    double rcsq = rcut * rcut;
    long np = 0;
    for (int i = 0; i < n; i++)
        for (int j = i + 1; j < n; j++) {
            double dx[3];
            int sh[3];
            min_image(L, xyz + 3 * i, xyz + 3 * j, dx, sh);
            double rsq = dx[0] * dx[0] + dx[1] * dx[1] + dx[2] * dx[2];
            if (rsq <= rcsq) {
                if (np < cap) {
                    int32_t *o = out + 5 * np;
                    o[0] = i; o[1] = j; o[2] = sh[0]; o[3] = sh[1]; o[4] = sh[2];
                }
                np++;
            }
        }
    return np;
}

/* Link-cell restatement (neighbours.F90:356-1027 criterion r^2 <= rcut^2,
 * cells of edge >= rcut, periodic images instead of halo copies).  Same
 * output convention as orc_pairs_brute.  Falls back to brute force when a
 * dimension has fewer than 3 cells. */
long orc_pairs_cells(int n, const double *cell, const double *xyz, double rcut, int32_t *out, long cap)
{
    double L[3] = {cell[0], cell[4], cell[8]};
    int nc[3];
    for (int d = 0; d < 3; d++) nc[d] = (int)(L[d] / rcut);
    if (nc[0] < 3 || nc[1] < 3 || nc[2] < 3) return orc_pairs_brute(n, cell, xyz, rcut, out, cap);
    int ncell = nc[0] * nc[1] * nc[2];
    int *head = malloc(sizeof(int) * ncell), *next = malloc(sizeof(int) * n), *cid = malloc(sizeof(int) * n);
    for (int c = 0; c < ncell; c++) head[c] = -1;
    for (int i = n - 1; i >= 0; i--) {
        int c3[3];
        for (int d = 0; d < 3; d++) {
            double s = xyz[3 * i + d] / L[d] + 0.5;
            s -= floor(s);
            int k = (int)(s * nc[d]);
            if (k >= nc[d]) k = nc[d] - 1;
            if (k < 0) k = 0;
            c3[d] = k;
        }
        cid[i] = (c3[2] * nc[1] + c3[1]) * nc[0] + c3[0];
        next[i] = head[cid[i]];
        head[cid[i]] = i;
    }
    double rcsq = rcut * rcut;
    long np = 0;
    for (int i = 0; i < n; i++) {
        int c = cid[i];
        int cx = c % nc[0], cy = (c / nc[0]) % nc[1], cz = c / (nc[0] * nc[1]);
        for (int oz = -1; oz <= 1; oz++)
            for (int oy = -1; oy <= 1; oy++)
                for (int ox = -1; ox <= 1; ox++) {
                    int jx = (cx + ox + nc[0]) % nc[0], jy = (cy + oy + nc[1]) % nc[1], jz = (cz + oz + nc[2]) % nc[2];
                    for (int j = head[(jz * nc[1] + jy) * nc[0] + jx]; j >= 0; j = next[j]) {
                        if (j <= i) continue;
                        double dx[3];
                        int sh[3];
                        min_image(L, xyz + 3 * i, xyz + 3 * j, dx, sh);
                        double rsq = dx[0] * dx[0] + dx[1] * dx[1] + dx[2] * dx[2];
                        if (rsq <= rcsq) {
                            if (np < cap) {
                                int32_t *o = out + 5 * np;
                                o[0] = i; o[1] = j; o[2] = sh[0]; o[3] = sh[1]; o[4] = sh[2];
                            }
                            np++;
                        }
                    }
                }
    }
    free(head); free(next); free(cid);
    return np;
}

/* ------------------------------------------------------------------ */
/* B-splines: bspline.F90:192-306 (values, 1st and 2nd derivatives)    */
/* ------------------------------------------------------------------ */
/* derivs[(l*S + t)] for l = 0,1,2 and t = 0..S-1 (t = reference index j-1);
 * point t sits at 0-based grid index floor(u) - S + 1 + t. */
void orc_bspline(double u, int S, double *derivs)
{
    double ncombk[22][22];
    memset(ncombk, 0, sizeof(ncombk));
    for (int nn = 1; nn <= S; nn++)
        for (int k = 0; k <= nn; k++) {
            double pr = 1.0;
            for (int i = nn - k + 1; i <= nn; i++) pr *= (double)i;
            int kk = k > 1 ? k : 1;
            double pi_ = 1.0;
            for (int i = 1; i <= kk; i++) pi_ *= 1.0 / (double)i;
            ncombk[k][nn] = pr * pi_;
        }
    const int nd = 2;
    double zd[24], fd[24];
    for (int i = 0; i < 24; i++) zd[i] = 0.0;
    zd[S] = u - trunc(u);
    zd[S - 1] = 1.0 - zd[S];
    double centre = zd[S];
    for (int k = S - 2; k >= nd + 1; k--) {
        double k_r = (double)(S - k + 1);
        double km1_rr = 1.0 / (double)(S - k);
        for (int j = k; j <= S - 1; j++) {
            double pt = centre + (double)(S - j);
            zd[j] = (pt * zd[j] + (k_r - pt) * zd[j + 1]) * km1_rr;
        }
        zd[S] = zd[S] * centre * km1_rr;
    }
    for (int l = nd; l >= 1; l--) {
        double k_r = (double)(S - l + 1);
        double km1_rr = 1.0 / (double)(S - l);
        for (int i = 0; i < 24; i++) fd[i] = 0.0;
        for (int j = 1; j <= S - 1; j++) {
            int lim = l < S - j ? l : S - j;
            for (int n = 0; n <= lim; n += 2) {
                fd[j] = fd[j] + ncombk[n][l] * zd[j + n];
                fd[j] = fd[j] - ncombk[n + 1][l] * zd[j + n + 1];
            }
            double pt = centre + (double)(S - j);
            zd[j] = (pt * zd[j] + (k_r - pt) * zd[j + 1]) * km1_rr;
        }
        for (int j = 1; j <= S - 1; j++) derivs[l * S + (j - 1)] = fd[j];
        derivs[l * S + (S - 1)] = zd[S];
        zd[S] = zd[S] * centre * km1_rr;
    }
    for (int j = 1; j <= S; j++) derivs[j - 1] = zd[j];
}

/* bspline.F90:73-190 (+ spl_cexp 308-353): norm2[m] = |b(m)|^2, m = 0..K-1 */
void orc_bspline_norm2(int K, int S, double *norm2)
{
    const double twopi = 2.0 * ORC_PI;
    double *wr = malloc(sizeof(double) * K), *wi = malloc(sizeof(double) * K);
    wr[0] = 1.0; wi[0] = 0.0;
    for (int i = 1; i <= K / 2; i++) {
        double arg = (twopi / (double)K) * (double)i;
        wr[i] = cos(arg); wi[i] = sin(arg);
        wr[K - i] = wr[i]; wi[K - i] = -wi[i];
    }
    double cs[24];
    for (int i = 0; i < 24; i++) cs[i] = 0.0;
    cs[2] = 1.0;
    for (int k = 3; k <= S; k++)
        for (int j = k; j >= 2; j--)
            cs[j] = ((double)(j - 1) * cs[j] + (double)(k - j + 1) * cs[j - 1]) * (1.0 / (double)(k - 1));
    for (int i = 0; i < K; i++) {
        double tr = 0.0, ti = 0.0;
        for (int k = 0; k <= S - 2; k++) {
            int m = (int)(((long)i * k) % K);
            tr += cs[k + 2] * wr[m];
            ti += cs[k + 2] * wi[m];
        }
        int m = (int)(((long)i * (S - 1)) % K);
        /* b = w / t */
        double den = tr * tr + ti * ti;
        double br = (wr[m] * tr + wi[m] * ti) / den;
        double bi = (wi[m] * tr - wr[m] * ti) / den;
        norm2[i] = br * br + bi * bi;
    }
    free(wr); free(wi);
}

/* ------------------------------------------------------------------ */
/* FFT: complex, unnormalised, sign +1 => sum_k c_k exp(+2 pi i j k/n)  */
/* (gpfa235.F90:72 convention).  Recursive mixed radix for 2^a 3^b 5^c; */
/* any other length falls back to a direct DFT.                         */
/* ------------------------------------------------------------------ */
static void fft_rec(const double *in, int stride, double *out, int n, int sign, double *tmp)
{
    if (n == 1) { out[0] = in[0]; out[1] = in[1]; return; }
    int p = (n % 2 == 0) ? 2 : (n % 3 == 0) ? 3 : (n % 5 == 0) ? 5 : n;
    int m = n / p;
    if (p == n) { /* direct DFT */
        for (int k = 0; k < n; k++) {
            double sr = 0.0, si = 0.0;
            for (int j = 0; j < n; j++) {
                double a = sign * 2.0 * ORC_PI * (double)(((long)j * k) % n) / (double)n;
                double c = cos(a), s = sin(a);
                double xr = in[2 * j * stride], xi = in[2 * j * stride + 1];
                sr += xr * c - xi * s; si += xr * s + xi * c;
            }
            out[2 * k] = sr; out[2 * k + 1] = si;
        }
        return;
    }
    /* decimation in time: sub-transform r takes elements r, r+p, r+2p, ... */
    for (int r = 0; r < p; r++) fft_rec(in + 2 * r * stride, stride * p, tmp + 2 * r * m, m, sign, out);
    for (int k = 0; k < m; k++) {
        for (int s = 0; s < p; s++) {
            double sr = 0.0, si = 0.0;
            int kk = k + s * m;
            for (int r = 0; r < p; r++) {
                long e = ((long)r * kk) % n;
                double a = sign * 2.0 * ORC_PI * (double)e / (double)n;
                double c = cos(a), sn = sin(a);
                double xr = tmp[2 * (r * m + k)], xi = tmp[2 * (r * m + k) + 1];
                sr += xr * c - xi * sn; si += xr * sn + xi * c;
            }
            out[2 * kk] = sr; out[2 * kk + 1] = si;
        }
    }
}

void orc_fft1d(double *data, int n, int sign)
{
    double *a = malloc(sizeof(double) * 2 * n), *b = malloc(sizeof(double) * 2 * n), *t = malloc(sizeof(double) * 2 * n);
    memcpy(a, data, sizeof(double) * 2 * n);
    fft_rec(a, 1, b, n, sign, t);
    memcpy(data, b, sizeof(double) * 2 * n);
    free(a); free(b); free(t);
}

/* grid[(z*K2 + y)*K1 + x] interleaved complex; transforms x, y then z */
void orc_fft3d(double *grid, const int *K, int sign)
{
    int K1 = K[0], K2 = K[1], K3 = K[2];
    int mx = K1 > K2 ? K1 : K2; if (K3 > mx) mx = K3;
    double *line = malloc(sizeof(double) * 2 * mx);
    for (int z = 0; z < K3; z++)
        for (int y = 0; y < K2; y++) {
            double *p = grid + 2 * ((long)(z * K2 + y) * K1);
            orc_fft1d(p, K1, sign);
        }
    for (int z = 0; z < K3; z++)
        for (int x = 0; x < K1; x++) {
            for (int y = 0; y < K2; y++) { long o = 2 * ((long)(z * K2 + y) * K1 + x); line[2 * y] = grid[o]; line[2 * y + 1] = grid[o + 1]; }
            orc_fft1d(line, K2, sign);
            for (int y = 0; y < K2; y++) { long o = 2 * ((long)(z * K2 + y) * K1 + x); grid[o] = line[2 * y]; grid[o + 1] = line[2 * y + 1]; }
        }
    for (int y = 0; y < K2; y++)
        for (int x = 0; x < K1; x++) {
            for (int z = 0; z < K3; z++) { long o = 2 * ((long)(z * K2 + y) * K1 + x); line[2 * z] = grid[o]; line[2 * z + 1] = grid[o + 1]; }
            orc_fft1d(line, K3, sign);
            for (int z = 0; z < K3; z++) { long o = 2 * ((long)(z * K2 + y) * K1 + x); grid[o] = line[2 * z]; grid[o + 1] = line[2 * z + 1]; }
        }
    free(line);
}

/* ------------------------------------------------------------------ */
/* the two-body force evaluation                                       */
/* ------------------------------------------------------------------ */
typedef struct {
    int natms, ntypes;
    const double *cell;      /* 9, row vectors */
    const double *xyz;       /* N*3 */
    const double *chge;      /* N */
    const int32_t *ltype;    /* N, 0-based */
    double rcut, padding, rvdw, alpha, epsq;
    int kvec[3], nspl, electro;
    int max_grid;            /* vdW tables: nkey rows of max_grid+1 */
    const double *vdw_pot, *vdw_frc;
    int nsamples;            /* erfc tables: nsamples values, tab[i-1] = f(i*dr) */
    const double *erfc_t, *erfc_d;
    double elrc, vlrc;
    int vdw_fs;              /* vdW force-shifting (vdws%l_force_shift) */
    int vdw_direct;          /* vdw_method direct: analytic forms per pair (vdws%l_direct) */
    const int *vdw_kind;     /* per key, for vdw_direct */
    const double *vdw_param; /* 7 per key, for vdw_direct */
    const int32_t *excl_off; /* excluded pairs (neigh%list_excl) as CSR over atoms: N+1 row starts, */
    const int32_t *excl_ids; /* then 0-based partners, both ends listed; NULL = none */
    const int32_t *frozen;   /* per atom, nonzero = frozen (config%lfrzn); NULL = none */
} orc_input;

typedef struct {
    double *frc;             /* N*3 out (overwritten) */
    double buffer[20];       /* two_body.F90:708-727 slots */
    double stress[9];        /* two_body_forces contribution to stats%stress */
    double engcpe, vircpe, engsrp, virsrp; /* two_body.F90:766-782 increments */
    double eng_self, eng_recip, eng_nz;
    double recip_net[3];     /* mean reciprocal force removed (ewald_general.F90:857-864) */
    long npairs;             /* pairs within rcut+padding */
    /* optional intermediates (NULL to skip): K1*K2*K3 reals each */
    double *charge_grid, *potential_grid;
    /* per-particle energies (N) and stresses (9N, column-major 3x3 per atom)
     * of stats%collect_pp (NULL to skip): overwritten */
    double *pp_energy, *pp_stress;
    double engvdw_abs;       /* sum of |vdW pair energy|: the scale of engsrp's rounding */
    double virvdw_abs;       /* sum of |vdW pair virial|: the scale of virsrp's rounding */
} orc_output;

static int excluded(const orc_input *in, int i, int j)
{
    if (!in->excl_off) return 0;
    for (int e = in->excl_off[i]; e < in->excl_off[i + 1]; e++)
        if (in->excl_ids[e] == j) return 1;
    return 0;
}

/* ewald_spole.F90:479-679 for one excluded pair at separation r = rrr <
 * rcut, charge product chgprd (with the 1/(4 pi eps0 eps) scaling):
 * the smooth erf(alpha r)/r part the reciprocal sum holds, taken back.  erf
 * is A&S 7.1.26, or its Taylor series to x^9 below alpha r = 0.01. */
void orc_excl_pair(double chgprd, double alpha, double rrr, double *erfr_out, double *egamma_out)
{
    const double a1 = 0.254829592, a2 = -0.284496736, a3 = 1.421413741, a4 = -1.453152027, a5 = 1.061405429,
                 pp = 0.3275911, sqrpi = sqrt(ORC_PI);
    double rsq = rrr * rrr, alpr = rrr * alpha, alpr2 = alpr * alpr, erfr, egamma;
    if (alpr < 1.0e-2) {
        erfr = 2.0 * chgprd * (alpha / sqrpi) *
               (1.0 + alpr2 * (-1.0 / 3.0 + alpr2 * (0.1 + alpr2 * (-1.0 / 42.0 + alpr2 / 216.0))));
        egamma = -4.0 * chgprd * (alpha * alpha * alpha / sqrpi) *
                 (1.0 / 3.0 + alpr2 * (-0.2 + alpr2 * (3.0 / 42.0 - 4.0 * alpr2 / 216.0)));
    } else {
        double e = exp(-(alpha * rrr) * (alpha * rrr));
        double tt = 1.0 / (1.0 + pp * alpha * rrr);
        erfr = chgprd * (1.0 - tt * (a1 + tt * (a2 + tt * (a3 + tt * (a4 + tt * a5)))) * e) / rrr;
        egamma = -(erfr - 2.0 * chgprd * (alpha / sqrpi) * e) / rsq;
    }
    *erfr_out = erfr;
    *egamma_out = egamma;
}

/* per-particle share of one pair i < j (vdw.F90:1987-1997, ewald_spole.F90:
 * 205-213): half the energy and half of r (x) f (statistics.F90:2616-2625:
 * element a + 3b = r_a f_b, f the force on i, r = x_i - x_j) to each atom.
 * eng_j is the energy share j receives: the tabulated vdW loop evaluates the
 * pair energy only on the side that passes the halo tie-break (vdw.F90:
 * 1946-1966, ltg(i) < ltg(j) for a pair across the cell boundary), so there
 * the other atom receives none. */
static void pp_pair(orc_output *out, int i, int j, double eng, double eng_j, double xx, double yy, double zz,
                    double fx, double fy, double fz)
{
    if (out->pp_energy) {
        out->pp_energy[i] += eng * 0.5;
        out->pp_energy[j] += eng_j * 0.5;
    }
    if (out->pp_stress) {
        const double r[3] = {xx, yy, zz}, f[3] = {fx, fy, fz};
        for (int b = 0; b < 3; b++)
            for (int a = 0; a < 3; a++) {
                out->pp_stress[9 * i + a + 3 * b] += r[a] * f[b] * 0.5;
                out->pp_stress[9 * j + a + 3 * b] += r[a] * f[b] * 0.5;
            }
    }
}

static int key_of(int ti, int tj)
{
    int a = ti + 1, b = tj + 1, hi = a > b ? a : b, lo = a > b ? b : a;
    return hi * (hi - 1) / 2 + lo - 1;
}

int orc_two_body(const orc_input *in, orc_output *out)
{
    const int n = in->natms;
    const double *c = in->cell;
    if (!(c[1] == 0.0 && c[2] == 0.0 && c[3] == 0.0 && c[5] == 0.0 && c[6] == 0.0 && c[7] == 0.0)) return 2;
    double L[3] = {c[0], c[4], c[8]};
    double *f = out->frc;
    memset(f, 0, sizeof(double) * 3 * n);
    memset(out->buffer, 0, sizeof(out->buffer));
    memset(out->stress, 0, sizeof(out->stress));
    double scaling = R4PIE0 / in->epsq;     /* two_body.F90:188 */
    double volm = fabs(L[0] * L[1] * L[2]);
    double *ppe = out->pp_energy, *pps = out->pp_stress;
    if (ppe) memset(ppe, 0, sizeof(double) * n);
    if (pps) memset(pps, 0, sizeof(double) * 9 * n);

    /* ---- reciprocal space: ewald_spole.F90:244-477 ---- */
    double engcpe_rc = 0.0, vircpe_rc = 0.0, sumchg = 0.0, self = 0.0;
    for (int i = 0; i < n; i++) sumchg += in->chge[i];
    int any_q = 0;
    for (int i = 0; i < n; i++) if (fabs(in->chge[i]) > ZERO_PLUS) any_q = 1;
    if (in->electro && any_q) {
        const int S = in->nspl, K1 = in->kvec[0], K2 = in->kvec[1], K3 = in->kvec[2];
        const long KK = (long)K1 * K2 * K3;
        /* spme.F90:159-231 self interaction (pot_order 1: alpha * 1/sqrt(pi)) */
        double sq = 0.0;
        for (int i = 0; i < n; i++) sq += in->chge[i] * in->chge[i];
        self = -sq * (scaling * in->alpha * (1.0 / sqrt(ORC_PI)) / 1.0);
        double rcell[9], det;
        orc_invert(c, rcell, &det);
        if (fabs(det) < 1.0e-6) return 120;
        double kr[3] = {(double)K1, (double)K2, (double)K3};
        double scale = ORC_PI * sqrt(ORC_PI) * pow(in->alpha, 1 - 3) * (0.5 * (1.0 / volm)) * scaling;
        int *idx = malloc(sizeof(int) * 3 * n);
        double *spl = malloc(sizeof(double) * 3 * 3 * S * n);
        for (int i = 0; i < n; i++) {
            const double *r = in->xyz + 3 * i;
            for (int d = 0; d < 3; d++) {
                double u = kr[d] * (rcell[d] * r[0] + rcell[d + 3] * r[1] + rcell[d + 6] * r[2] + 0.5);
                /* An atom that drifted out of the cell between list rebuilds is
                 * represented by its periodic image inside [0, K): the reference
                 * spreads that image from its halo (ewald_spole.F90:381-384). */
                if (u < 0.0 || u >= kr[d]) u -= kr[d] * floor(u / kr[d]);
                idx[3 * i + d] = (int)u;
                orc_bspline(u, S, spl + (3 * i + d) * 3 * S);
            }
        }
        double *Q = calloc(KK, sizeof(double));
        /* ewald_general.F90:517-576, periodic images folded with (g mod K) */
        for (int i = 0; i < n; i++) {
            double qi = in->chge[i];
            if (fabs(qi) <= ZERO_PLUS) continue;
            const double *Mx = spl + (3 * i + 0) * 3 * S, *My = spl + (3 * i + 1) * 3 * S, *Mz = spl + (3 * i + 2) * 3 * S;
            for (int tz = 0; tz < S; tz++) {
                int gz = ((idx[3 * i + 2] - S + 1 + tz) % K3 + K3) % K3;
                double f1 = qi * Mz[tz];
                for (int ty = 0; ty < S; ty++) {
                    int gy = ((idx[3 * i + 1] - S + 1 + ty) % K2 + K2) % K2;
                    double f2 = f1 * My[ty];
                    for (int tx = 0; tx < S; tx++) {
                        int gx = ((idx[3 * i] - S + 1 + tx) % K1 + K1) % K1;
                        Q[((long)gz * K2 + gy) * K1 + gx] += f2 * Mx[tx];
                    }
                }
            }
        }
        if (out->charge_grid) memcpy(out->charge_grid, Q, sizeof(double) * KK);
        /* ewald_spole.F90:1257-1386 */
        double *G = malloc(sizeof(double) * 2 * KK);
        for (long g = 0; g < KK; g++) { G[2 * g] = Q[g]; G[2 * g + 1] = 0.0; }
        orc_fft3d(G, in->kvec, +1);
        double *Gs = NULL;  /* FFT+(Q) kept for the per-particle stress grid */
        double pst[6] = {0, 0, 0, 0, 0, 0};  /* sum of the per-particle stresses (unscaled) */
        if (pps) {
            Gs = malloc(sizeof(double) * 2 * KK);
            memcpy(Gs, G, sizeof(double) * 2 * KK);
        }
        double *n2x = malloc(sizeof(double) * K1), *n2y = malloc(sizeof(double) * K2), *n2z = malloc(sizeof(double) * K3);
        orc_bspline_norm2(K1, S, n2x); orc_bspline_norm2(K2, S, n2y); orc_bspline_norm2(K3, S, n2z);
        double rcp[10];
        orc_dcell(rcell, rcp);
        double recip_conv_fac = ORC_PI / in->alpha;
        double test_fac = pow(1.0e-6 / recip_conv_fac, 2);
        double m1 = kr[0] * rcp[6], m2 = kr[1] * rcp[7], m3 = kr[2] * rcp[8];
        double mn = m1 < m2 ? m1 : m2; if (m3 < mn) mn = m3;
        double cut_off = 0.5 * 1.05 * mn, cut_off_2 = cut_off * cut_off;
        const double sqrpi = sqrt(ORC_PI);
        double st[3][3] = {{0}};
        for (int l = 0; l < K3; l++) {
            int ll = l; if (2 * ll > K3) ll -= K3;
            double rp3[3] = {ll * rcell[2], ll * rcell[5], ll * rcell[8]};
            double bb3 = n2z[l];
            for (int k = 0; k < K2; k++) {
                int kk = k; if (2 * kk > K2) kk -= K2;
                double rp2[3] = {rp3[0] + kk * rcell[1], rp3[1] + kk * rcell[4], rp3[2] + kk * rcell[7]};
                double bb2 = bb3 * n2y[k];
                for (int j = 0; j < K1; j++) {
                    int jj = j; if (2 * jj > K1) jj -= K1;
                    double rp[3] = {rp2[0] + jj * rcell[0], rp2[1] + jj * rcell[3], rp2[2] + jj * rcell[6]};
                    double bb1 = bb2 * n2x[j];
                    double k2 = rp[0] * rp[0] + rp[1] * rp[1] + rp[2] * rp[2];
                    long o = 2 * (((long)l * K2 + k) * K1 + j);
                    if (k2 <= cut_off_2 && k2 > test_fac) {
                        double fp = sqrt(k2) * recip_conv_fac;
                        double fac = bb1 * exp(-(fp * fp)) / (sqrpi * (fp * fp));
                        double pr = fac * G[o], pi_ = fac * G[o + 1];
                        double w = -2.0 * ((1.0 + fp * fp) / k2);
                        /* Real(pc * w * conj(Q)) */
                        double pv = (pr * G[o] + pi_ * G[o + 1]) * w;
                        for (int a = 0; a < 3; a++)
                            for (int b = 0; b < 3; b++) st[a][b] += rp[a] * rp[b] * pv;
                        G[o] = pr; G[o + 1] = pi_;
                    } else {
                        G[o] = 0.0; G[o + 1] = 0.0;
                    }
                }
            }
        }
        orc_fft3d(G, in->kvec, -1);
        if (out->potential_grid) for (long g = 0; g < KK; g++) out->potential_grid[g] = G[2 * g];
        /* ewald_general.F90:717-869 */
        double rk[3];
        /* Matmul(reshape(rcell,[3,3]), K): element d = sum_e rcell(d + 3(e-1)) K_e */
        for (int d = 0; d < 3; d++) rk[d] = rcell[d] * kr[0] + rcell[d + 3] * kr[1] + rcell[d + 6] * kr[2];
        double eng = 0.0, ftot[3] = {0, 0, 0};
        double *fr = calloc(3 * n, sizeof(double));
        for (int i = 0; i < n; i++) {
            double qi = in->chge[i];
            if (fabs(qi) <= ZERO_PLUS) continue;
            const double *Sx = spl + (3 * i + 0) * 3 * S, *Sy = spl + (3 * i + 1) * 3 * S, *Sz = spl + (3 * i + 2) * 3 * S;
            double et = 0.0, cf[3] = {0, 0, 0};
            for (int tz = 0; tz < S; tz++) {
                int gz = ((idx[3 * i + 2] - S + 1 + tz) % K3 + K3) % K3;
                double e2 = qi * Sz[tz];
                double f13 = qi * Sz[tz], f23 = qi * Sz[tz], f33 = qi * Sz[S + tz];
                for (int ty = 0; ty < S; ty++) {
                    int gy = ((idx[3 * i + 1] - S + 1 + ty) % K2 + K2) % K2;
                    double e1 = e2 * Sy[ty];
                    double f12 = f13 * Sy[ty], f22 = f23 * Sy[S + ty], f32 = f33 * Sy[ty];
                    for (int tx = 0; tx < S; tx++) {
                        int gx = ((idx[3 * i] - S + 1 + tx) % K1 + K1) % K1;
                        double phi = G[2 * (((long)gz * K2 + gy) * K1 + gx)];
                        double fx = f12 * Sx[S + tx] * phi * rk[0];
                        double fy = f22 * Sx[tx] * phi * rk[1];
                        double fz = f32 * Sx[tx] * phi * rk[2];
                        ftot[0] -= fx; ftot[1] -= fy; ftot[2] -= fz;
                        cf[0] += fx; cf[1] += fy; cf[2] += fz;
                        et += e1 * Sx[tx] * phi;
                    }
                }
            }
            eng += et;
            if (ppe) ppe[i] += et * scale;  /* ewald_general.F90:852, ewald_spole.F90:429,465 */
            fr[3 * i] -= cf[0]; fr[3 * i + 1] -= cf[1]; fr[3 * i + 2] -= cf[2];
        }
        if (pps) {
            /* per-particle reciprocal stress (ewald_general.F90:578-715 with
             * stress_kernel :1033-1057, then spme_calc_stress :871-1009): the
             * charge transform times B(m) f'(x) pi/(x alpha), x = pi|m|/alpha,
             * f(x) = exp(-x^2)/(sqrt(pi) x^2) (spme.F90:233-247,421-436),
             * transformed back, then gathered with second derivatives of the
             * splines */
            for (int l = 0; l < K3; l++) {
                int ll = l; if (2 * ll > K3) ll -= K3;
                for (int k = 0; k < K2; k++) {
                    int kk = k; if (2 * kk > K2) kk -= K2;
                    for (int j = 0; j < K1; j++) {
                        int jj = j; if (2 * jj > K1) jj -= K1;
                        /* the same summation order as the potential loop above */
                        double r3[3] = {ll * rcell[2], ll * rcell[5], ll * rcell[8]};
                        double r2[3] = {r3[0] + kk * rcell[1], r3[1] + kk * rcell[4], r3[2] + kk * rcell[7]};
                        double rp[3] = {r2[0] + jj * rcell[0], r2[1] + jj * rcell[3], r2[2] + jj * rcell[6]};
                        double bb1 = n2z[l] * n2y[k] * n2x[j];
                        double k2 = rp[0] * rp[0] + rp[1] * rp[1] + rp[2] * rp[2];
                        long o = 2 * (((long)l * K2 + k) * K1 + j);
                        double x = ORC_PI * sqrt(k2) * (1.0 / in->alpha);
                        if (k2 <= cut_off_2 && x > 1.0e-6) {
                            double energy = exp(-(x * x)) / (sqrpi * (x * x));
                            double fpd = (-2.0 / x * energy) - ((2.0 / x) * (1.0 / sqrpi) * exp(-(x * x)));
                            double fac = fpd * ORC_PI / x / in->alpha;
                            Gs[o] = bb1 * Gs[o] * fac;
                            Gs[o + 1] = bb1 * Gs[o + 1] * fac;
                        } else {
                            Gs[o] = 0.0; Gs[o + 1] = 0.0;
                        }
                    }
                }
            }
            orc_fft3d(Gs, in->kvec, -1);
            const double c_fac = 1.0 / (2.0 * 2.0 * ORC_PI * in->alpha);
            for (int m = 0; m < 6; m++) pst[m] = 0.0;
            for (int i = 0; i < n; i++) {
                double qi = in->chge[i];
                const double *Sx = spl + (3 * i + 0) * 3 * S, *Sy = spl + (3 * i + 1) * 3 * S, *Sz = spl + (3 * i + 2) * 3 * S;
                /* lower triangle (1,1) (2,1) (3,1) (2,2) (3,2) (3,3) */
                double st6[6] = {0, 0, 0, 0, 0, 0};
                for (int tz = 0; tz < S; tz++) {
                    int gz = ((idx[3 * i + 2] - S + 1 + tz) % K3 + K3) % K3;
                    double z0 = qi * Sz[tz], z1 = qi * Sz[S + tz], z2 = qi * Sz[2 * S + tz];
                    for (int ty = 0; ty < S; ty++) {
                        int gy = ((idx[3 * i + 1] - S + 1 + ty) % K2 + K2) % K2;
                        double y0 = Sy[ty], y1 = Sy[S + ty], y2 = Sy[2 * S + ty];
                        double t11 = z0 * y0, t21 = z0 * y1, t31 = z1 * y0, t22 = z0 * y2, t32 = z1 * y1, t33 = z2 * y0;
                        for (int tx = 0; tx < S; tx++) {
                            int gx = ((idx[3 * i] - S + 1 + tx) % K1 + K1) % K1;
                            double g = Gs[2 * (((long)gz * K2 + gy) * K1 + gx)];
                            double x0 = Sx[tx], x1 = Sx[S + tx], x2 = Sx[2 * S + tx];
                            double v[6] = {t11 * x2 * g, t21 * x1 * g, t31 * x1 * g, t22 * x0 * g, t32 * x0 * g, t33 * x0 * g};
                            const int ab[6][2] = {{0, 0}, {1, 0}, {2, 0}, {1, 1}, {2, 1}, {2, 2}};
                            for (int m = 0; m < 6; m++)
                                st6[m] -= v[m] * rk[ab[m][0]] * rk[ab[m][1]] * c_fac;
                        }
                    }
                }
                const int map9[9] = {0, 1, 2, 1, 3, 4, 2, 4, 5};
                for (int k = 0; k < 9; k++) pps[9 * i + k] += st6[map9[k]] * scale;
                for (int m = 0; m < 6; m++) pst[m] += st6[m];
            }
            free(Gs);
        }
        if (ppe) for (int i = 0; i < n; i++) ppe[i] += self / (double)n;  /* ewald_spole.F90:465 */
        for (int d = 0; d < 3; d++) { ftot[d] /= (double)n; out->recip_net[d] = ftot[d]; }
        for (int i = 0; i < n; i++)
            for (int d = 0; d < 3; d++) fr[3 * i + d] -= ftot[d];
        eng *= scale;
        for (int i = 0; i < 3 * n; i++) f[i] += fr[i] * scale * 2.0;
        double s9[9];
        /* stress_temp(beta,alpha) column-major reshape => s9[(alpha-1)*3 + beta-1] */
        for (int a = 0; a < 3; a++)
            for (int b = 0; b < 3; b++) s9[a * 3 + b] = st[a][b] * scale;
        if (pps) {
            /* with per-particle collection the reference's reciprocal stress
             * is the sum of the per-particle ones (ewald_general.F90:963,1002
             * overwrite the k-space sum) */
            const int map9[9] = {0, 1, 2, 1, 3, 4, 2, 4, 5};
            for (int k = 0; k < 9; k++) s9[k] = pst[map9[k]] * scale;
        }
        s9[0] += eng; s9[4] += eng; s9[8] += eng;
        for (int k = 0; k < 9; k++) out->stress[k] += s9[k];
        vircpe_rc = -(s9[0] + s9[4] + s9[8]);
        engcpe_rc = eng + self;
        out->eng_recip = eng;
        free(fr); free(G); free(Q); free(idx); free(spl); free(n2x); free(n2y); free(n2z);
    }
    out->eng_self = self;

    /* ---- real space: two_body.F90:339-525, vdw.F90:1790-2024, ewald_spole.F90:58-242 ---- */
    double rcut_ext = in->rcut + in->padding;
    long cap = 0;
    long np = orc_pairs_cells(n, c, in->xyz, rcut_ext, NULL, 0);
    cap = np;
    int32_t *pairs = malloc(sizeof(int32_t) * 5 * (cap > 0 ? cap : 1));
    orc_pairs_cells(n, c, in->xyz, rcut_ext, pairs, cap);
    out->npairs = np;
    double engvdw = 0.0, virvdw = 0.0, engrl = 0.0, virrl = 0.0, engvdw_abs = 0.0, virvdw_abs = 0.0;
    double st_v[6] = {0}, st_c[6] = {0};
    double dlrpot = in->rvdw / (double)(in->max_grid - 4), rdr = 1.0 / dlrpot;
    double dr_e = in->rcut / (double)(in->nsamples - 4), rdr_e = 1.0 / dr_e;
    int nrow = in->max_grid + 1;
    long nexcl_listed = 0;
    for (long p = 0; p < np; p++) {
        int i = pairs[5 * p], j = pairs[5 * p + 1];
        /* neighbours.F90:1198-1250: frozen-frozen and excluded pairs sit past
         * list(0,i), out of the two-body loop */
        if ((in->frozen && in->frozen[i] && in->frozen[j]) || excluded(in, i, j)) { nexcl_listed++; continue; }
        double xx = in->xyz[3 * i] - (in->xyz[3 * j] + pairs[5 * p + 2] * L[0]);
        double yy = in->xyz[3 * i + 1] - (in->xyz[3 * j + 1] + pairs[5 * p + 3] * L[1]);
        double zz = in->xyz[3 * i + 2] - (in->xyz[3 * j + 2] + pairs[5 * p + 4] * L[2]);
        double rrr = sqrt(xx * xx + yy * yy + zz * zz);
        /* vdW, direct: vdw.F90:1655-1700 with vdw_direct_fs_generate (:969-1048) */
        if (in->vdw_direct && rrr < in->rvdw) {
            int k = key_of(in->ltype[i], in->ltype[j]);
            int kind = in->vdw_kind[k];
            if (kind > 0) {
                double pk[7], afs = 0.0, bfs = 0.0, e, g;
                orc_vdw_params(kind, in->vdw_param + 7 * (long)k, in->rvdw, pk);
                /* vdw.F90:1015-1040: none for n-m shift, DPD, nDPD, LJF, 12-6-MDF */
                if (in->vdw_fs && kind != 7 && kind != 10 && kind != 23 && kind != 21 && kind != 20) {
                    double z, dz;
                    orc_vdw_eval(kind, pk, in->rvdw, &z, &dz);
                    afs = dz / in->rvdw;
                    bfs = -z - dz;
                }
                double r_rrr = 1.0 / rrr, rsq = rrr * rrr, r_rsq = r_rrr * r_rrr;
                orc_vdw_eval(kind, pk, rrr, &e, &g);
                double eng = e + afs * rrr + bfs;
                double gamma = g * r_rsq - afs * r_rrr;
                double fx = gamma * xx, fy = gamma * yy, fz = gamma * zz;
                f[3 * i] += fx; f[3 * i + 1] += fy; f[3 * i + 2] += fz;
                f[3 * j] -= fx; f[3 * j + 1] -= fy; f[3 * j + 2] -= fz;
                engvdw += eng;
                engvdw_abs += fabs(eng);
                virvdw -= gamma * rsq;
                virvdw_abs += fabs(gamma * rsq);
                st_v[0] += xx * fx; st_v[1] += xx * fy; st_v[2] += xx * fz;
                st_v[3] += yy * fy; st_v[4] += yy * fz; st_v[5] += zz * fz;
                pp_pair(out, i, j, eng, eng, xx, yy, zz, fx, fy, fz);
            }
        }
        /* vdW, tabulated */
        if (!in->vdw_direct && in->max_grid > 0 && rrr < in->rvdw) {
            int k = key_of(in->ltype[i], in->ltype[j]);
            const double *tf = in->vdw_frc + (long)k * nrow, *tp = in->vdw_pot + (long)k * nrow;
            double r_rrr = 1.0 / rrr, rsq = rrr * rrr, r_rsq = r_rrr * r_rrr;
            int l = (int)(rrr * rdr);
            double ppp = rrr * rdr - (double)l;
            double gk = tf[l]; if (l == 0) gk = gk * rrr;
            double gk1 = tf[l + 1], gk2 = tf[l + 2];
            double t1 = gk + (gk1 - gk) * ppp, t2 = gk1 + (gk2 - gk1) * (ppp - 1.0);
            double gamma = (t1 + (t2 - t1) * ppp * 0.5) * r_rsq;
            /* vdw.F90:1922: force-shifting with the samples at r = rvdw (index max_grid-4) */
            double r_rvdw = 1.0 / in->rvdw;
            if (in->vdw_fs) gamma = gamma - tf[in->max_grid - 4] * (r_rrr * r_rvdw);
            double fx = gamma * xx, fy = gamma * yy, fz = gamma * zz;
            f[3 * i] += fx; f[3 * i + 1] += fy; f[3 * i + 2] += fz;
            f[3 * j] -= fx; f[3 * j + 1] -= fy; f[3 * j + 2] -= fz;
            double vk = tp[l], vk1 = tp[l + 1], vk2 = tp[l + 2];
            t1 = vk + (vk1 - vk) * ppp; t2 = vk1 + (vk2 - vk1) * (ppp - 1.0);
            double eng = t1 + (t2 - t1) * ppp * 0.5;
            /* vdw.F90:1961-1964 */
            if (in->vdw_fs) eng = eng + tf[in->max_grid - 4] * (rrr * r_rvdw - 1.0) - tp[in->max_grid - 4];
            engvdw += eng;
            engvdw_abs += fabs(eng);
            virvdw -= gamma * rsq;
            virvdw_abs += fabs(gamma * rsq);
            st_v[0] += xx * fx; st_v[1] += xx * fy; st_v[2] += xx * fz;
            st_v[3] += yy * fy; st_v[4] += yy * fz; st_v[5] += zz * fz;
            const int cross = pairs[5 * p + 2] != 0 || pairs[5 * p + 3] != 0 || pairs[5 * p + 4] != 0;
            pp_pair(out, i, j, eng, cross ? 0.0 : eng, xx, yy, zz, fx, fy, fz);
        }
        /* erfc-damped Coulomb */
        if (in->electro) {
            double qi = in->chge[i] * scaling, qj = in->chge[j];
            if (fabs(qi) >= ZERO_PLUS && fabs(qj) > ZERO_PLUS && rrr < in->rcut) {
                double prefac = qi * qj;
                double eg = prefac * orc_interp3(in->erfc_d - 1, rdr_e, rrr);
                double fx = eg * xx, fy = eg * yy, fz = eg * zz;
                f[3 * i] += fx; f[3 * i + 1] += fy; f[3 * i + 2] += fz;
                f[3 * j] -= fx; f[3 * j + 1] -= fy; f[3 * j + 2] -= fz;
                double ec = prefac * orc_interp3(in->erfc_t - 1, rdr_e, rrr);
                engrl += ec;
                virrl -= eg * (rrr * rrr);
                pp_pair(out, i, j, ec, ec, xx, yy, zz, fx, fy, fz);
                st_c[0] += xx * fx; st_c[1] += xx * fy; st_c[2] += xx * fz;
                st_c[3] += yy * fy; st_c[4] += yy * fz; st_c[5] += zz * fz;
            }
        }
    }
    free(pairs);
    out->npairs = np - nexcl_listed;
    const int map[9] = {0, 1, 2, 1, 3, 4, 2, 4, 5};
    for (int k = 0; k < 9; k++) out->stress[k] += st_v[map[k]] + st_c[map[k]];

    /* ---- Ewald exclusion correction: two_body.F90:551-592, ewald_spole.F90:479-679 ----
     * every excluded pair once (i < j), at its minimum image, if both charges
     * are nonzero and r < rcut */
    double engcpe_ex = 0.0, vircpe_ex = 0.0, st_x[6] = {0};
    if (in->electro && in->excl_off) {
        for (int i = 0; i < n; i++) {
            double qi = in->chge[i];
            if (!(fabs(qi) > ZERO_PLUS)) continue;
            for (int e = in->excl_off[i]; e < in->excl_off[i + 1]; e++) {
                int j = in->excl_ids[e];
                if (j <= i) continue;
                double qj = in->chge[j];
                double dx[3];
                int sh[3];
                min_image(L, in->xyz + 3 * i, in->xyz + 3 * j, dx, sh);
                double rrr = sqrt(dx[0] * dx[0] + dx[1] * dx[1] + dx[2] * dx[2]);
                if (!(fabs(qj) > ZERO_PLUS && rrr < in->rcut)) continue;
                double erfr, egamma;
                orc_excl_pair(qj * (qi * scaling), in->alpha, rrr, &erfr, &egamma);
                double fx = egamma * dx[0], fy = egamma * dx[1], fz = egamma * dx[2];
                f[3 * i] += fx; f[3 * i + 1] += fy; f[3 * i + 2] += fz;
                f[3 * j] -= fx; f[3 * j + 1] -= fy; f[3 * j + 2] -= fz;
                engcpe_ex -= erfr;
                vircpe_ex -= egamma * rrr * rrr;
                st_x[0] += dx[0] * fx; st_x[1] += dx[0] * fy; st_x[2] += dx[0] * fz;
                st_x[3] += dx[1] * fy; st_x[4] += dx[1] * fz; st_x[5] += dx[2] * fz;
            }
        }
        for (int k = 0; k < 9; k++) out->stress[k] += st_x[map[k]];
    }

    /* ---- bookkeeping: two_body.F90:689-790 ---- */
    double engcpe_nz = 0.0, vircpe_nz = 0.0;
    if (in->electro && fabs(sumchg) > 1.0e-6) {
        double factor_nz = -0.5 * (ORC_PI * R4PIE0 / in->epsq) * pow(sumchg / in->alpha, 2);
        engcpe_nz = factor_nz / volm;
        vircpe_nz = -3.0 * engcpe_nz;
    }
    out->eng_nz = engcpe_nz;
    out->buffer[7] = engvdw; out->buffer[8] = virvdw;
    out->engvdw_abs = engvdw_abs;
    out->virvdw_abs = virvdw_abs;
    out->buffer[9] = engcpe_rc; out->buffer[10] = vircpe_rc;
    out->buffer[11] = engrl; out->buffer[12] = virrl;
    out->buffer[15] = engcpe_ex; out->buffer[16] = vircpe_ex;
    out->engcpe = engcpe_rc + engrl + engcpe_ex + engcpe_nz;
    out->vircpe = vircpe_rc + virrl + vircpe_ex + vircpe_nz;
    double tmp = -vircpe_nz / 3.0;
    out->stress[0] += tmp; out->stress[4] += tmp; out->stress[8] += tmp;
    out->engsrp = engvdw + in->elrc;
    out->virsrp = virvdw + in->vlrc;
    if (ppe) for (int i = 0; i < n; i++) ppe[i] += in->elrc / (double)n;  /* two_body.F90:783 */
    tmp = -in->vlrc / 3.0;
    out->stress[0] += tmp; out->stress[4] += tmp; out->stress[8] += tmp;
    return 0;
}

/* nve.F90:127-193 (stage 1) and 194-215 (stage 2) for one species-free
 * velocity-Verlet step; mass <= 1e-6 atoms are left untouched. */
void orc_nve_first(int n, double dt, const double *mass, double *x, double *v, const double *f)
{
    for (int i = 0; i < n; i++) {
        if (mass[i] <= 1.0e-6) continue;
        double tmp = (0.5 * dt) / mass[i];
        for (int d = 0; d < 3; d++) {
            v[3 * i + d] += tmp * f[3 * i + d];
            x[3 * i + d] += dt * v[3 * i + d];
        }
    }
}

void orc_nve_second(int n, double dt, const double *mass, double *v, const double *f)
{
    for (int i = 0; i < n; i++) {
        if (mass[i] <= 1.0e-6) continue;
        double tmp = (0.5 * dt) / mass[i];
        for (int d = 0; d < 3; d++) v[3 * i + d] += tmp * f[3 * i + d];
    }
}
