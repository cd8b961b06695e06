"""ctypes front end of the C restatement (``dlpx_oracle.c``) -- TEST INFRASTRUCTURE ONLY.

Imported by ``tests/``, ``__graft_entry__.smoke()`` and nothing in the
product.  ``build()`` compiles ``liboracle.so`` next to this file with gcc.
"""
import ctypes as C
import math
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "dlpx_oracle.c")
LIB = os.path.join(HERE, "liboracle.so")
DELR_MAX = 0.01  # constants.F90:139


def build():
    cmd = ["gcc", "-O2", "-fPIC", "-shared", "-std=c11", "-Wall", "-o", LIB, SRC, "-lm"]
    subprocess.run(cmd, check=True)
    return LIB


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB) or os.path.getmtime(LIB) < os.path.getmtime(SRC):
            build()
        _lib = C.CDLL(LIB)
        _lib.orc_erfc_as.restype = C.c_double
        _lib.orc_interp3.restype = C.c_double
        _lib.orc_pairs_brute.restype = C.c_long
        _lib.orc_pairs_cells.restype = C.c_long
    return _lib


def _p(a, t=C.c_double):
    return a.ctypes.data_as(C.POINTER(t))


class OrcInput(C.Structure):
    _fields_ = [("natms", C.c_int), ("ntypes", C.c_int),
                ("cell", C.POINTER(C.c_double)), ("xyz", C.POINTER(C.c_double)),
                ("chge", C.POINTER(C.c_double)), ("ltype", C.POINTER(C.c_int32)),
                ("rcut", C.c_double), ("padding", C.c_double), ("rvdw", C.c_double),
                ("alpha", C.c_double), ("epsq", C.c_double),
                ("kvec", C.c_int * 3), ("nspl", C.c_int), ("electro", C.c_int),
                ("max_grid", C.c_int), ("vdw_pot", C.POINTER(C.c_double)),
                ("vdw_frc", C.POINTER(C.c_double)),
                ("nsamples", C.c_int), ("erfc_t", C.POINTER(C.c_double)),
                ("erfc_d", C.POINTER(C.c_double)),
                ("elrc", C.c_double), ("vlrc", C.c_double), ("vdw_fs", C.c_int), ("vdw_direct", C.c_int),
                ("vdw_kind", C.POINTER(C.c_int)), ("vdw_param", C.POINTER(C.c_double)),
                ("excl_off", C.POINTER(C.c_int32)), ("excl_ids", C.POINTER(C.c_int32)),
                ("frozen", C.POINTER(C.c_int32))]


class OrcOutput(C.Structure):
    _fields_ = [("frc", C.POINTER(C.c_double)), ("buffer", C.c_double * 20),
                ("stress", C.c_double * 9), ("engcpe", C.c_double), ("vircpe", C.c_double),
                ("engsrp", C.c_double), ("virsrp", C.c_double), ("eng_self", C.c_double),
                ("eng_recip", C.c_double), ("eng_nz", C.c_double), ("recip_net", C.c_double * 3),
                ("npairs", C.c_long), ("charge_grid", C.POINTER(C.c_double)),
                ("potential_grid", C.POINTER(C.c_double)), ("pp_energy", C.POINTER(C.c_double)),
                ("pp_stress", C.POINTER(C.c_double)), ("engvdw_abs", C.c_double),
                ("virvdw_abs", C.c_double)]


def nint(x):
    return int(math.floor(x + 0.5)) if x >= 0 else -int(math.floor(-x + 0.5))


def table_size(cut):
    """bounds.F90:811,820: max(1004, nint(cut/delr_max) + 4)."""
    return max(1004, nint(cut / DELR_MAX) + 4)


def vdw_tables(sysm, max_grid=None):
    """Per pair key rows of length max_grid+1 (vdw.F90:1397-1576)."""
    mg = max_grid or table_size(sysm.vdw_cutoff)
    nkey = sysm.ntypes * (sysm.ntypes + 1) // 2
    pot = np.zeros((nkey, mg + 1)); frc = np.zeros((nkey, mg + 1))
    from_key = {}
    for (ti, tj, kind, params) in sysm.vdw:
        a, b = ti + 1, tj + 1
        hi, lo = max(a, b), min(a, b)
        k = hi * (hi - 1) // 2 + lo - 1
        p = np.zeros(7); p[:len(params)] = params
        pp = np.ascontiguousarray(p)
        rp = np.zeros(mg + 1); rf = np.zeros(mg + 1)
        if lib().orc_vdw_table(C.c_int(kind), _p(pp), C.c_double(sysm.vdw_cutoff), C.c_int(mg), _p(rp), _p(rf)):
            raise ValueError("unsupported vdW kind %d" % kind)
        pot[k] = rp; frc[k] = rf
        from_key[k] = (kind, p)
    return mg, pot, frc, from_key


def erfc_tables(sysm):
    n = table_size(sysm.cutoff)
    et = np.zeros(n); ed = np.zeros(n)
    lib().orc_erfc_tables(C.c_double(sysm.cutoff), C.c_double(sysm.alpha), C.c_int(n), _p(et), _p(ed))
    return n, et, ed


def vdw_lrc(sysm, from_key):
    nkey = sysm.ntypes * (sysm.ntypes + 1) // 2
    kinds = np.zeros(nkey, np.int32); params = np.zeros((nkey, 7))
    for k, (kind, p) in from_key.items():
        kinds[k] = kind; params[k] = p
    num = np.bincount(sysm.types, minlength=sysm.ntypes).astype(np.float64)
    frz = getattr(sysm, "frozen", None)
    nfz = None
    if frz is not None and np.any(frz):
        nfz = np.bincount(sysm.types[np.asarray(frz, bool)], minlength=sysm.ntypes).astype(np.float64)
    e = C.c_double(); v = C.c_double()
    lib().orc_vdw_lrc(C.c_int(sysm.ntypes), _p(num), _p(kinds, C.c_int), _p(params),
                      C.c_double(sysm.vdw_cutoff), C.c_double(sysm.volume), C.byref(e), C.byref(v),
                      _p(nfz) if nfz is not None else None)
    return e.value, v.value


def erfc_tables_exact(sysm, n=None):
    """Tables of the true erfc(alpha r)/r and its exact derivative term (for
    checks that need a force table consistent with the energy table)."""
    from scipy import special
    n = n or table_size(sysm.cutoff)
    r = np.arange(1, n + 1) * (sysm.cutoff / (n - 4))
    a = sysm.alpha
    et = special.erfc(a * r) / r
    ed = (et + a * 2.0 / math.sqrt(math.pi) * np.exp(-(a * r) ** 2)) / r ** 2
    return n, et, ed


def two_body(sysm, pos=None, grids=False, exact_erfc=False, table_points=None, per_particle=False):
    """One call of the restated two_body_forces (with its own pair search).
    exact_erfc=True swaps the reference's A&S erfc tables for the true erfc;
    table_points overrides the number of table samples (finer tables make the
    interpolated force the derivative of the interpolated energy to O(dr^2)).

    Returns dict: frc (N,3), buffer (20), stress (9), engcpe, vircpe, engsrp,
    virsrp, eng_self, eng_recip, npairs and, with grids=True, the real charge
    grid Q and potential grid Re(Phi) shaped (K3, K2, K1); with
    per_particle=True, pp_energy (N) and pp_stress (N, 9; element a + 3b of
    atom i is stats%pp_stress(a+3b+1, i))."""
    L = lib()
    mg, pot, frc, from_key = vdw_tables(sysm, table_points)
    ns, et, ed = erfc_tables_exact(sysm, table_points) if exact_erfc else erfc_tables(sysm)
    fs = bool(getattr(sysm, "vdw_force_shift", False))
    direct = bool(getattr(sysm, "vdw_direct", False))
    nkey = sysm.ntypes * (sysm.ntypes + 1) // 2
    kinds = np.zeros(nkey, np.int32); params = np.zeros((nkey, 7))
    for k, (kind, p) in from_key.items():
        kinds[k] = kind; params[k] = p
    # vdw.F90:662: no long-range correction with force-shifting
    elrc, vlrc = (0.0, 0.0) if fs else vdw_lrc(sysm, from_key)
    xyz = np.ascontiguousarray(sysm.pos if pos is None else pos, dtype=np.float64)
    cell = np.ascontiguousarray(sysm.cell, dtype=np.float64)
    q = np.ascontiguousarray(sysm.q)
    lt = np.ascontiguousarray(sysm.types, dtype=np.int32)
    pot = np.ascontiguousarray(pot); frc = np.ascontiguousarray(frc)
    inp = OrcInput(natms=sysm.natms, ntypes=sysm.ntypes, cell=_p(cell), xyz=_p(xyz), chge=_p(q),
                   ltype=_p(lt, C.c_int32), rcut=sysm.cutoff, padding=sysm.padding,
                   rvdw=sysm.vdw_cutoff, alpha=sysm.alpha, epsq=sysm.epsilon,
                   kvec=(C.c_int * 3)(*sysm.kvec), nspl=sysm.nsplines,
                   electro=1 if sysm.electrostatics else 0, max_grid=mg, vdw_pot=_p(pot),
                   vdw_frc=_p(frc), nsamples=ns, erfc_t=_p(et), erfc_d=_p(ed), elrc=elrc, vlrc=vlrc,
                   vdw_fs=1 if fs else 0, vdw_direct=1 if direct else 0, vdw_kind=_p(kinds, C.c_int),
                   vdw_param=_p(params))
    excl = sysm.exclusions() if hasattr(sysm, "exclusions") else None
    if excl is not None:
        eo = np.zeros(sysm.natms + 1, np.int32)
        eo[1:] = np.cumsum([len(e) for e in excl])
        ei = np.ascontiguousarray(np.concatenate(excl) if eo[-1] else np.zeros(1), dtype=np.int32)
        inp.excl_off = _p(eo, C.c_int32)
        inp.excl_ids = _p(ei, C.c_int32)
    frz = getattr(sysm, "frozen", None)
    if frz is not None and np.any(frz):
        fz = np.ascontiguousarray(np.asarray(frz, bool).astype(np.int32))
        inp.frozen = _p(fz, C.c_int32)
    f = np.zeros((sysm.natms, 3))
    out = OrcOutput(frc=_p(f))
    K = sysm.kvec
    if grids and sysm.electrostatics:
        Q = np.zeros(K[0] * K[1] * K[2]); P = np.zeros_like(Q)
        out.charge_grid = _p(Q); out.potential_grid = _p(P)
    if per_particle:
        ppe = np.zeros(sysm.natms); pps = np.zeros((sysm.natms, 9))
        out.pp_energy = _p(ppe); out.pp_stress = _p(pps)
    rc = L.orc_two_body(C.byref(inp), C.byref(out))
    if rc != 0:
        raise RuntimeError("orc_two_body returned %d" % rc)
    res = dict(frc=f, buffer=np.array(out.buffer[:]), stress=np.array(out.stress[:]),
               engcpe=out.engcpe, vircpe=out.vircpe, engsrp=out.engsrp, virsrp=out.virsrp,
               eng_self=out.eng_self, eng_recip=out.eng_recip, eng_nz=out.eng_nz,
               recip_net=np.array(out.recip_net[:]), npairs=out.npairs, elrc=elrc, vlrc=vlrc,
               engvdw_abs=out.engvdw_abs, virvdw_abs=out.virvdw_abs)
    if grids and sysm.electrostatics:
        res["Q"] = Q.reshape(K[2], K[1], K[0]); res["Phi"] = P.reshape(K[2], K[1], K[0])
    if per_particle:
        res["pp_energy"] = ppe; res["pp_stress"] = pps
    return res


def excl_pair(chgprd, alpha, r):
    """ewald_spole.F90:479-679 for one excluded pair: (erfr, egamma), with
    erfr = chgprd erf(alpha r)/r and egamma its force factor."""
    e, g = C.c_double(), C.c_double()
    lib().orc_excl_pair(C.c_double(chgprd), C.c_double(alpha), C.c_double(r), C.byref(e), C.byref(g))
    return e.value, g.value


def pairs(cell, xyz, rcut, method="cells"):
    """Unordered pair set within rcut: int32 array (npairs, 5) = (i, j, sx, sy, sz), i < j."""
    L = lib()
    fn = L.orc_pairs_cells if method == "cells" else L.orc_pairs_brute
    cell = np.ascontiguousarray(cell, np.float64); xyz = np.ascontiguousarray(xyz, np.float64)
    n = xyz.shape[0]
    cnt = fn(C.c_int(n), _p(cell), _p(xyz), C.c_double(rcut), None, C.c_long(0))
    out = np.zeros((max(cnt, 1), 5), np.int32)
    fn(C.c_int(n), _p(cell), _p(xyz), C.c_double(rcut), _p(out, C.c_int32), C.c_long(cnt))
    return out[:cnt]


def bspline(u, S=8):
    """(3, S) array: values, first and second derivatives at the S points below u."""
    d = np.zeros(3 * S)
    lib().orc_bspline(C.c_double(u), C.c_int(S), _p(d))
    return d.reshape(3, S)


def bspline_norm2(K, S=8):
    out = np.zeros(K)
    lib().orc_bspline_norm2(C.c_int(K), C.c_int(S), _p(out))
    return out


def fft3d(grid, sign):
    """grid complex128 shaped (K3, K2, K1); returns the unnormalised transform."""
    g = np.ascontiguousarray(grid, dtype=np.complex128).copy()
    K = (C.c_int * 3)(g.shape[2], g.shape[1], g.shape[0])
    lib().orc_fft3d(g.ctypes.data_as(C.POINTER(C.c_double)), K, C.c_int(sign))
    return g


def nve(sysm, steps, pos=None, vel=None):
    """``steps`` velocity-Verlet NVE steps (nve.F90:127-215) driven by two_body;
    returns (pos, vel, frc, last two_body result)."""
    L = lib()
    x = np.ascontiguousarray(sysm.pos if pos is None else pos, dtype=np.float64).copy()
    v = np.ascontiguousarray(sysm.vel if vel is None else vel, dtype=np.float64).copy()
    m = np.ascontiguousarray(sysm.mass)
    frz = getattr(sysm, "frozen", None)
    frz = np.asarray(frz, bool) if frz is not None else np.zeros(sysm.natms, bool)
    if vel is None:
        # set_temperature (temperature.F90:579-620, getvom kinetics.F90:365-419):
        # frozen atoms at rest, the free atoms' centre-of-mass velocity removed
        v[frz] = 0.0
        mf = m[~frz]
        if mf.sum() > 0.0:
            vom = (mf[:, None] * v[~frz]).sum(0) / mf.sum()
            v[~frz & (m > 1.0e-6)] -= vom

    def forces():
        # freeze_atoms (configuration.F90:3240-3258) after each evaluation
        r = two_body(sysm, pos=x)
        fr = r["frc"].copy()
        fr[frz] = 0.0
        v[frz] = 0.0
        return r, fr

    res, f = forces()
    for _ in range(steps):
        L.orc_nve_first(C.c_int(sysm.natms), C.c_double(sysm.timestep), _p(m), _p(x), _p(v), _p(f))
        res, f = forces()
        L.orc_nve_second(C.c_int(sysm.natms), C.c_double(sysm.timestep), _p(m), _p(v), _p(f))
    return x, v, f, res
