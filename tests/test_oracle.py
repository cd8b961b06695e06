"""The oracle (C restatement in oracle/dlpx_oracle.c) pinned against the
compiled reference's golden fixtures and against what the mathematics fixes.
CPU only."""
import math
import os

import numpy as np
import pytest
from scipy import integrate, special

import fixtures
import oracle
from dlpx import fileio, systems

STATIC = [n for n in fixtures.NAMES if "nve" not in n]
TRAJ = [n for n in fixtures.NAMES if "nve" in n]


def _force_margin(f, ref):
    """|f - ref| in units of half the last printed digit of REVCON (g20.10),
    with a floor of 1e-12 max|F| for components that are nearly zero."""
    bound = fixtures.printed_digits_bound(ref) + 1e-12 * np.abs(ref).max()
    return float((np.abs(f - ref) / bound).max())


@pytest.mark.parametrize("name", STATIC)
def test_oracle_matches_reference_fixture(name):
    s, ref = fixtures.load(name)
    r = oracle.two_body(s)
    # forces: every printed digit of the reference (10 significant figures)
    assert _force_margin(fixtures.frozen_zeroed(s, r["frc"]), ref["frc"]) <= 1.01
    # energies and virials from the binary REVIVE record (full precision),
    # relative to the largest term of the sum: with exclusions engcpe is a
    # small difference of the self, reciprocal and correction energies, and
    # engsrp can be a small sum of large pair energies of both signs (the
    # ZBL forms' repulsion against attractive pairs); 1e-14 of the summed
    # magnitudes covers the rounding of a different summation order
    b = r["buffer"]
    comps = dict(engcpe=(r["eng_self"], r["eng_recip"], b[11], b[15]), vircpe=(b[10], b[12], b[16]),
                 engsrp=(1e-2 * r["engvdw_abs"],))
    for k in ("engsrp", "engcpe", "virsrp", "vircpe"):
        scale = max([abs(ref[k]), 1.0] + [abs(x) for x in comps.get(k, ())])
        assert abs(r[k] - ref[k]) <= 1e-12 * scale, k
    assert np.abs(r["stress"] - ref["stress"]).max() <= 1e-12 * np.abs(ref["stress"]).max()


@pytest.mark.parametrize("name", STATIC)
def test_oracle_matches_reference_full_precision(name):
    """The reference's own step-0 forces, dumped at full double precision
    after its two_body_forces call (tests/golden/ref_dump.F90), not only
    REVCON's 10 printed digits: RMS relative 1e-12, every component within
    1e-11 of the largest; the dumped energies against the oracle's."""
    s, ref = fixtures.load(name)
    r = oracle.two_body(s)
    f, f64 = r["frc"], ref["frc64"]
    assert float(np.sqrt(((f - f64) ** 2).sum() / (f64 ** 2).sum())) <= 1e-12
    assert np.abs(f - f64).max() <= 1e-11 * np.abs(f64).max()
    b = r["buffer"]
    comps = dict(engcpe=(r["eng_self"], r["eng_recip"], b[11], b[15]), vircpe=(b[10], b[12], b[16]),
                 engsrp=(1e-2 * r["engvdw_abs"],))
    for k in ("engsrp", "engcpe", "virsrp", "vircpe"):
        scale = max([abs(ref[k]), 1.0] + [abs(x) for x in comps.get(k, ())])
        assert abs(r[k] - ref["dump_" + k]) <= 1e-12 * scale, k
    assert np.abs(r["stress"] - ref["dump_stress"]).max() <= 1e-12 * np.abs(ref["dump_stress"]).max()


@pytest.mark.parametrize("name", [n for n in STATIC if "tab_pot" in fixtures.load(n)[1]])
def test_oracle_vdw_tables_match_reference_tables(name):
    """Every sample of the oracle's generated vdW tables against the
    reference's own vdws%tab_potential / tab_force (vdw.F90:1397-1576) for
    the forms of the fixture: equal to a few ulps (sample 0 holds Huge in
    both), and the long-range corrections against vdws%elrc / vlrc."""
    s, ref = fixtures.load(name)
    mg, pot, frc, from_key = oracle.vdw_tables(s)
    rmg, rpot, rfrc, ltp = fixtures.reference_tables(s, ref)
    assert mg == rmg
    for key in from_key:
        for mine, theirs in ((pot[key], rpot[key]), (frc[key], rfrc[key])):
            # rounding of each sample (powers of up to r^-14), with a floor of 1e-13 of the table's
            # scale for samples where terms cancel (a zero crossing, the
            # shifted n-m form's correction)
            floor = 1e-13 * np.abs(theirs[len(theirs) // 10:]).max()
            d = np.abs(mine[1:] - theirs[1:])
            assert (d <= 1e-14 * np.abs(theirs[1:]) + floor).all(), (key, from_key[key][0], d.max())
    if not s.vdw_force_shift:
        # 1e-11: the AMOEBA correction is a Simpson sum of over 10^4 terms
        elrc, vlrc = oracle.vdw_lrc(s, from_key)
        assert abs(elrc - ref["elrc"]) <= 1e-11 * max(abs(ref["elrc"]), 1e-300)
        assert abs(vlrc - ref["vlrc"]) <= 1e-11 * max(abs(ref["vlrc"]), 1e-300)


@pytest.mark.parametrize("name", TRAJ)
def test_oracle_nve_matches_reference_trajectory(name):
    s, ref = fixtures.load(name)
    x, v, f, r = oracle.nve(s, ref["steps"])
    L = np.diag(s.cell.reshape(3, 3))
    dx = x - ref["pos"]
    dx -= L * np.round(dx / L)  # the reference folds positions at list rebuilds
    assert (np.abs(dx) / (fixtures.printed_digits_bound(ref["pos"]) + 1e-12)).max() <= 1.01
    assert (np.abs(v - ref["vel"]) / (fixtures.printed_digits_bound(ref["vel"]) + 1e-12)).max() <= 1.01
    assert _force_margin(f, ref["frc"]) <= 1.01
    for k in ("engsrp", "engcpe"):
        assert abs(r[k] - ref[k]) <= 1e-11 * abs(ref[k]), k


def _eval(kind, p, r, rvdw=None):
    """U and gamma of one form; with rvdw, the parameters are first completed
    as potential_cutoffs does (the damped forms' outer radius)."""
    import ctypes as C
    e, g = C.c_double(), C.c_double()
    pp = np.zeros(7)
    pp[:len(p)] = p
    if rvdw is not None:
        q = np.zeros(7)
        oracle.lib().orc_vdw_params(C.c_int(kind), pp.ctypes.data_as(C.POINTER(C.c_double)), C.c_double(rvdw),
                                    q.ctypes.data_as(C.POINTER(C.c_double)))
        pp = q
    oracle.lib().orc_vdw_eval(C.c_int(kind), pp.ctypes.data_as(C.POINTER(C.c_double)), C.c_double(r),
                              C.byref(e), C.byref(g))
    return e.value, g.value


# test_vdw.F90:46-59: energy and virial (-gamma) of each form at r = 1 A with
# parameters 1, 2, ..., 7 and a 10 A cutoff (index = vdW key)
VDW_KNOWN = {1: (-1.0, 0.0), 2: (16128.0, -195072.0), 3: (80.0, -288.0),
             4: (-2.3934693402873668, 17.696734670143684), 5: (45.598150033144236, -45.196300066288472),
             6: (-1.0, 8.0), 8: (363.25771964635982, -2300.0595394172847), 9: (1.0, 12.0), 10: (0.25, -0.5),
             12: (15616.0, -192000.0), 13: (367.25771964635982, -2348.0595394172847),
             14: (2.1495939317213679, 0.71653131057378916), 22: (-0.8948393168143698, -0.1988531815143044),
             23: (-0.4921875, -0.09375), 7: (4567.8016528926664, -16899.173553719396),
             11: (555.50928442334305, -3719.0014773362891),
             15: (16882.652704957309, -50044.066263312896), 16: (3.8142266913481757E+030, -5.2445617006037446E+031),
             17: (11761.890149027380, -36135.103419546060), 18: (16128.0, -195072.0),
             19: (-2.3934693402873668, 17.696734670143684), 20: (-1.0, 0.0), 21: (4.0, -40.0),
             24: (-1765.997274794514, 15903.665444480088)}


@pytest.mark.parametrize("kind", sorted(VDW_KNOWN))
def test_vdw_known_answers(kind):
    e, g = _eval(kind, [1, 2, 3, 4, 5, 6, 7], 1.0, rvdw=10.0)
    ee, vv = VDW_KNOWN[kind]
    assert abs(e - ee) <= 1e-12 * max(abs(ee), 1.0) and abs(-g - vv) <= 1e-12 * max(abs(vv), 1.0)


@pytest.mark.parametrize("r", [2.1, 3.3, 7.9, 11.5])
def test_vdw_closed_forms_and_gamma(r):
    eps, sig = 99.6, 3.405
    e, g = _eval(2, [eps, sig], r)
    assert abs(e - 4 * eps * ((sig / r) ** 12 - (sig / r) ** 6)) <= 1e-12 * abs(e)
    A, B, s0, Cc, D = systems.TOSI_FUMI[(1, 0)]
    e, g = _eval(5, [A, B, s0, Cc, D], r)
    u = lambda x: A * math.exp(B * (s0 - x)) - Cc / x ** 6 - D / x ** 8
    assert abs(e - u(r)) <= 1e-12 * abs(e)
    h = 1e-5  # gamma = -r dU/dr
    assert abs(g - (-r * (u(r + h) - u(r - h)) / (2 * h))) <= 1e-6 * abs(g)


def test_erfc_tables_closed_form():
    rc, alpha, n = 12.0, 0.28, oracle.table_size(12.0)
    s = systems.nacl(ncells=5, kvec=(32, 32, 32))
    nn, et, ed = oracle.erfc_tables(s)
    assert nn == n == 1204
    r = np.arange(1, n + 1) * (rc / (n - 4))
    # A&S 7.1.26 stays within 1.5e-7 of the true erfc
    assert np.abs(et * r - special.erfc(alpha * r)).max() < 1.6e-7
    g = et
    dexact = (g + alpha * 2.0 / math.sqrt(math.pi) * np.exp(-(alpha * r) ** 2)) / r ** 2
    assert np.allclose(ed, dexact, rtol=1e-13, atol=0)


def test_interpolation_exact_for_quadratics():
    import ctypes as C
    a, b, c = 0.7, -1.3, 0.25
    dr = 0.01
    tab = np.array([a + b * (i * dr) + c * (i * dr) ** 2 for i in range(0, 1210)])
    p = tab.ctypes.data_as(C.POINTER(C.c_double))
    rng = np.random.default_rng(3)
    for r in rng.uniform(0.05, 11.9, 200):
        v = oracle.lib().orc_interp3(p, C.c_double(1.0 / dr), C.c_double(r))
        assert abs(v - (a + b * r + c * r * r)) < 1e-12


def _pairs_numpy(cell, x, rcut):
    L = np.diag(cell.reshape(3, 3))
    out = set()
    for i in range(len(x)):
        d = x[i] - x[i + 1:]
        s = np.rint(d / L)
        d = x[i] - (x[i + 1:] + s * L)
        r2 = (d * d).sum(1)
        for k in np.nonzero(r2 <= rcut * rcut)[0]:
            j = i + 1 + k
            out.add((i, j) + tuple(int(v) for v in s[k]))
    return out


@pytest.mark.parametrize("seed", [1, 2])
def test_pair_sets_cells_equal_bruteforce(seed):
    rng = np.random.default_rng(seed)
    L = np.array([25.0, 27.5, 31.0])
    x = rng.uniform(-0.5, 0.5, (400, 3)) * L
    cell = np.diag(L).reshape(9)
    a = {tuple(p) for p in oracle.pairs(cell, x, 8.2, "cells")}
    b = {tuple(p) for p in oracle.pairs(cell, x, 8.2, "brute")}
    c = _pairs_numpy(cell, x, 8.2)
    assert a == b == c and len(a) > 1000


def _bspline_closed(x, n):
    """Cardinal B-spline M_n(x) by the Cox-de Boor recursion (support [0, n])."""
    if n == 1:
        return 1.0 if 0.0 <= x < 1.0 else 0.0
    return (x * _bspline_closed(x, n - 1) + (n - x) * _bspline_closed(x - 1, n - 1)) / (n - 1)


@pytest.mark.parametrize("S", [4, 6, 8, 10, 12])
def test_bspline_values_and_derivatives(S):
    rng = np.random.default_rng(S)
    for u in rng.uniform(0, 50, 40):
        d = oracle.bspline(u, S)
        w = u - math.floor(u)
        # point t sits at floor(u) - S + 1 + t: value M_S(u - g) = M_S(w + S - 1 - t)
        closed = np.array([_bspline_closed(w + S - 1 - t, S) for t in range(S)])
        assert np.allclose(d[0], closed, atol=1e-14)
        assert abs(d[0].sum() - 1.0) < 1e-14          # partition of unity
        assert abs(d[1].sum()) < 1e-13               # derivatives sum to zero
        dclosed = np.array([_bspline_closed(w + S - 1 - t, S - 1) - _bspline_closed(w + S - 2 - t, S - 1)
                            for t in range(S)])
        assert np.allclose(d[1], dclosed, atol=1e-13)  # d/du M_S(u - g)
    # symmetry M_S(x) = M_S(S - x)
    a, b = oracle.bspline(10.3, S), oracle.bspline(10.7, S)
    assert np.allclose(a[0], b[0][::-1], atol=1e-14)


@pytest.mark.parametrize("K", [(6, 10, 15), (8, 9, 5), (32, 30, 36)])
def test_fft3d_convention(K):
    rng = np.random.default_rng(sum(K))
    g = rng.standard_normal((K[2], K[1], K[0])) + 1j * rng.standard_normal((K[2], K[1], K[0]))
    plus = oracle.fft3d(g, +1)
    minus = oracle.fft3d(g, -1)
    # gpfa235.F90:72: +1 is sum c_k exp(+2 pi i jk/n), unnormalised
    assert np.allclose(plus, np.fft.ifftn(g) * g.size, atol=1e-10)
    assert np.allclose(minus, np.fft.fftn(g), atol=1e-10)


@pytest.mark.parametrize("K,S", [(32, 8), (30, 6), (45, 8), (128, 8)])
def test_bspline_moduli(K, S):
    n2 = oracle.bspline_norm2(K, S)
    Mk = np.array([_bspline_closed(k + 1.0, S) for k in range(S - 1)])
    m = np.arange(K)
    den = (Mk[None, :] * np.exp(2j * np.pi * np.outer(m, np.arange(S - 1)) / K)).sum(1)
    b = np.exp(2j * np.pi * (S - 1) * m / K) / den
    assert np.allclose(n2, np.abs(b) ** 2, rtol=1e-12)


def test_madelung_rocksalt():
    s = systems.nacl(ncells=5, kvec=(64, 64, 64), jitter_on=False, temperature=0.0)
    r = oracle.two_body(s)
    r0 = s.meta["a"] / 2.0
    e_madelung = -s.natms / 2.0 * 1.747564594633182 * systems.R4PIE0 / r0
    # limited by the A&S erfc (|error| < 1.5e-7) in the real-space tables
    assert abs(r["engcpe"] - e_madelung) < 2e-6 * abs(e_madelung)
    assert np.abs(r["frc"]).max() < 1e-3  # perfect lattice: forces vanish by symmetry


def _ewald_recip(s, kcut):
    """Explicit reciprocal-space Ewald sum over m = (j/Lx, k/Ly, l/Lz), |m| <= kcut."""
    L = np.diag(s.cell.reshape(3, 3))
    V = L.prod()
    nmax = np.ceil(kcut * L).astype(int)
    mx, my, mz = np.meshgrid(*[np.arange(-n, n + 1) for n in nmax], indexing="ij")
    m = np.stack([mx.ravel() / L[0], my.ravel() / L[1], mz.ravel() / L[2]], 1)
    m2 = (m * m).sum(1)
    keep = (m2 > 0) & (m2 <= kcut * kcut)
    m, m2 = m[keep], m2[keep]
    q = s.q
    ph = np.exp(2j * np.pi * (s.pos @ m.T))
    S = (q[:, None] * ph).sum(0)
    w = np.exp(-(np.pi ** 2) * m2 / s.alpha ** 2) / m2
    E = systems.R4PIE0 / (2 * np.pi * V) * (w * np.abs(S) ** 2).sum()
    return E


def test_spme_energy_converges_to_explicit_ewald_sum():
    s = systems.nacl(ncells=3, kvec=(40, 40, 40), cutoff=8.5, padding=0.0, alpha=0.395, seed=5)
    r = oracle.two_body(s)
    L = s.cell[0]
    kcut = 0.5 * 1.05 * 40 / L  # the reference's spherical cutoff
    e = _ewald_recip(s, kcut)
    assert abs(r["eng_recip"] - e) < 1e-7 * abs(e)


def test_forces_are_minus_energy_gradient():
    # With the true erfc the real-space force table is the exact derivative of
    # the energy table (the reference's A&S erfc energy is not, at ~1e-4), and
    # with 100x finer tables the derivative of the 3-point interpolant of the
    # energy matches the interpolated force to ~1e-7 (at the reference's
    # 0.01 A spacing the O(dr^2) mismatch is ~1e-4 for the Born-Mayer wall).
    s = systems.nacl(ncells=3, kvec=(40, 40, 40), cutoff=8.5, padding=0.0, alpha=0.395, seed=9)
    fine = 85004
    r0 = oracle.two_body(s, exact_erfc=True, table_points=fine)
    rng = np.random.default_rng(0)
    h = 1e-5
    for i in rng.choice(s.natms, 4, replace=False):
        for d in range(3):
            xp = s.pos.copy(); xp[i, d] += h
            xm = s.pos.copy(); xm[i, d] -= h
            ep = oracle.two_body(s, pos=xp, exact_erfc=True, table_points=fine)
            em = oracle.two_body(s, pos=xm, exact_erfc=True, table_points=fine)
            fd = -((ep["engcpe"] + ep["engsrp"]) - (em["engcpe"] + em["engsrp"])) / (2 * h)
            # tabulated force and energy are separate 3-point interpolants and
            # the mean reciprocal force is removed
            assert abs(fd - r0["frc"][i, d]) < 1e-6 * np.abs(r0["frc"][i]).max() + 1e-4


def _tail(kind, p):
    """The part of U(r) whose tail the reference integrates (vdw.F90:697-782):
    the whole form, except Buckingham and BHM (dispersion terms only)."""
    if kind == 1:
        return lambda r: p[0] / r ** 12 - p[1] / r ** 6
    if kind == 2:
        return lambda r: 4 * p[0] * ((p[1] / r) ** 12 - (p[1] / r) ** 6)
    if kind == 3:
        return lambda r: p[0] / (p[1] - p[2]) * (p[2] * (p[3] / r) ** p[1] - p[1] * (p[3] / r) ** p[2])
    if kind == 4:
        return lambda r: -p[2] / r ** 6
    if kind == 5:
        return lambda r: -p[3] / r ** 6 - p[4] / r ** 8
    if kind == 6:
        return lambda r: p[0] / r ** 12 - p[1] / r ** 10
    if kind in (7, 9, 10, 22, 23):  # no correction for these forms (vdw.F90:681-933)
        return lambda r: 0.0
    if kind == 11:
        return lambda r: p[0] * (1.07 / (r / p[1] + 0.07)) ** 7 * (1.12 / ((r / p[1]) ** 7 + 0.12) - 2.0)
    if kind == 12:
        return lambda r: 4 * p[0] * ((p[1] / r) ** 12 - p[2] * (p[1] / r) ** 6)
    if kind == 13:
        return lambda r: p[0] * ((1 - math.exp(-p[2] * (r - p[1]))) ** 2 - 1) + p[3] / r ** 12
    if kind == 14:
        return lambda r: (p[0] + p[1] * r) * math.exp(-r / p[2])
    return lambda r: p[0] * ((1 - math.exp(-p[2] * (r - p[1]))) ** 2 - 1)


@pytest.mark.parametrize("make", [lambda: systems.nacl(ncells=5, kvec=(32, 32, 32)),
                                  lambda: systems.mixture(ncells=6),
                                  lambda: systems.mixture(ncells=6, vdw=systems.MIX2_VDW),
                                  lambda: systems.mixture(ncells=6, vdw=systems.MIX3_VDW)])
def test_vdw_lrc_matches_tail_integrals(make):
    s = make()
    _, _, _, from_key = oracle.vdw_tables(s)
    elrc, vlrc = oracle.vdw_lrc(s, from_key)
    V = s.volume
    n = np.bincount(s.types, minlength=s.ntypes).astype(float)
    rc = s.vdw_cutoff
    e = v = 0.0
    for (ti, tj, kind, p) in s.vdw:
        mult = 1.0 if ti == tj else 2.0
        fac = mult * 2 * np.pi * n[ti] * n[tj] / V
        if kind == 11:
            # AMOEBA: the reference's own quadrature (vdw.F90:522-613, called
            # at :795-796), whose chunks each recount one panel, so the energy
            # lies a few 1e-6 above the integral, and whose virial integrand is
            # not the derivative of the form
            e += fac * _amoeba_sum(p[1], p[0], rc, virial=False)
            v += fac * _amoeba_sum(p[1], p[0], rc, virial=True)
            continue
        ut = _tail(kind, p)
        h = 1e-4
        dut = lambda r: (ut(r + h) - ut(r - h)) / (2 * h)
        e += fac * integrate.quad(lambda r: r * r * ut(r), rc, np.inf)[0]
        v += fac * integrate.quad(lambda r: r ** 3 * dut(r), rc, np.inf)[0]
    assert abs(elrc - e) < 1e-9 * abs(e)
    assert abs(vlrc - v) < 1e-7 * abs(v)


def _amoeba_sum(r0, eps, rw, virial, a=0.07, b=0.12, prec=1.0e-12):
    """intRadMM3 / intRaddMM3 of vdw.F90:522-613 in numpy: Simpson panels of
    width h = (rw/r0)/10000 over chunks of 10000 steps from x = rw/r0, each
    chunk's first panel starting where the last one ended (so one panel per
    chunk is counted twice), until a chunk adds less than prec.  Returns the
    energy integral, or the virial one (the reference's padd with its sign
    flipped, which enters vlrc with the energy's sign)."""
    def mm3(x):
        return eps * ((1.0 + a) / (a + x)) ** 7 * ((1.0 + b) / (b + x ** 7) - 2.0)

    def dmm3(x):
        ia, ib = 1.0 / (a + x), 1.0 / (b + x ** 7)
        return -7.0 * (mm3(x) * ia - ((1.0 + a) * ia) ** 7 * ((1.0 + b) * ib ** 2) * x ** 6) / r0

    f, pw, sc = (dmm3, 3, 4) if virial else (mm3, 2, 3)
    n = 10000
    xs = rw / r0
    h = (2 * xs - xs) / n
    s, sold, j = 0.0, np.finfo(float).max, 1
    while abs(s - sold) * h / 3.0 > prec:
        sold = s
        x = xs + np.arange((j - 1) * n, j * n + 1, 2) * h
        s += float(np.sum(x ** pw * f(x) + 4.0 * (x + h) ** pw * f(x + h) + (x + 2.0 * h) ** pw * f(x + 2.0 * h)))
        j += 1
    return s * h * r0 ** sc / 3.0


def _erf_over_r_series(a, r, terms=30):
    """erf(a r)/r and (1/r) d/dr of it from the Maclaurin series of erf
    (alternating, fast for a r < 1)."""
    f = g = 0.0
    for k in range(terms):
        c = (-1) ** k * a ** (2 * k + 1) / (math.factorial(k) * (2 * k + 1))
        f += c * r ** (2 * k)
        if k:
            g += c * 2 * k * r ** (2 * k - 2)
    s = 2.0 / math.sqrt(math.pi)
    return s * f, s * g


@pytest.mark.parametrize("r", [1e-4, 0.004, 0.0285, 0.03, 0.5, 1.0, 2.7, 8.9])
def test_excl_pair_is_erf_over_r(r):
    """ewald_spole.F90:479-679 per pair: erfr = q erf(alpha r)/r and egamma =
    (1/r) d erfr/dr.  Against the Maclaurin series of erf (exact to rounding
    here): the series branch (alpha r < 0.01) to 1e-14, the A&S 7.1.26 branch
    within its published bound |erf error| <= 1.5e-7 (so 1.5e-7/r on erfr and,
    with the exact Gaussian term, 1.5e-7/r^3 on egamma)."""
    a, q = 0.35, 2.5e4
    e, g = oracle.excl_pair(q, a, r)
    fe, fg = _erf_over_r_series(a, r, terms=60)
    if a * r < 1e-2:
        assert e == pytest.approx(q * fe, rel=1e-14)
        assert g == pytest.approx(q * fg, rel=1e-12)
    else:
        assert abs(e - q * fe) <= q * 1.5e-7 / r
        assert abs(g - q * fg) <= q * 1.5e-7 / r ** 3


@pytest.mark.parametrize("name", STATIC)
def test_oracle_per_particle_matches_reference_heat_flux(name):
    """Per-particle energies and stresses (stats%collect_pp) against the
    reference's HEATFLUX record at step 0 (12 significant figures), which
    weights every atom's values with its own velocity, and the reference's
    stress and Coulomb virial of the same run, whose reciprocal part is the
    sum of the per-particle stresses (ewald_general.F90:963,1002)."""
    s, _ = fixtures.load(name)
    z = np.load(os.path.join(fixtures.GOLDEN, name + ".npz"))
    o = oracle.two_body(s, per_particle=True)
    want = z["ref_heatflux"]
    assert np.abs(fixtures.heat_flux(s, o["pp_energy"], o["pp_stress"]) - want).max() <= 1e-11 * np.abs(want).max()
    st = z["ref_pp_stress"]
    assert np.abs(o["stress"] - st).max() <= 1e-12 * np.abs(st).max()
    vc = z["ref_pp_stpval"][fileio.STP_VIRCPE]
    assert abs(o["vircpe"] - vc) <= 1e-12 * max(abs(vc), abs(o["buffer"][10]), abs(o["buffer"][16]), 1.0)


def test_oracle_per_particle_sums():
    """Sum rules of the per-particle split: with direct vdW (every pair's
    energy shared by both atoms) the energies add up to the two-body energy
    less the net-charge term, which has no per-particle part; without SPME the
    stresses add up to the two-body stress less the long-range-correction
    diagonal, which has none either."""
    s, _ = fixtures.load("mix_4000_direct")
    o = oracle.two_body(s, per_particle=True)
    total = o["engsrp"] + o["engcpe"] - o["eng_nz"]
    assert abs(o["pp_energy"].sum() - total) <= 1e-12 * abs(o["eng_self"])
    s, _ = fixtures.load("argon_4000_jit")
    o = oracle.two_body(s, per_particle=True)
    want = o["stress"].reshape(3, 3) + np.eye(3) * o["vlrc"] / 3.0
    got = o["pp_stress"].sum(0).reshape(3, 3).T
    assert np.abs(got - want).max() <= 1e-12 * np.abs(want).max()
