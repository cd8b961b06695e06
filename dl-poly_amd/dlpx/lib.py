"""ctypes binding of ``libdlpx.so`` (the C ABI declared in include/dlpx.h).

``Context`` mirrors the reference's call surface for the force path:
``link_cell_pairs`` (neighbours.F90:356), ``vnl_check`` (neighbours.F90:123)
and ``two_body_forces`` (two_body.F90:76), plus a resident MD loop
(``md_load`` / ``md_run`` / ``md_store``).  Arrays are numpy; positions are
(N, 3) float64 in DL_POLY internal units.  Errors raise ``DlpxError`` carrying
the DL_POLY error code (106, 95, 307, 120) or the library's own code.

The library must be built (``python -m dlpx.build``); there is no fallback.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "libdlpx.so")

FP64, MIXED = 0, 1
VDW_MAX_PARAM = 7  # parameters per vdW key (bounds.F90:666 vdws%max_param)

EXPORTS = ["dlpx_create", "dlpx_destroy", "dlpx_last_error", "dlpx_version", "dlpx_set_vdw_tables",
           "dlpx_set_erfc_tables", "dlpx_generate_vdw_tables", "dlpx_generate_erfc_tables", "dlpx_set_lrc",
           "dlpx_compute_lrc", "dlpx_get_lrc", "dlpx_link_cell_pairs", "dlpx_get_pairs", "dlpx_vnl_check",
           "dlpx_two_body_forces", "dlpx_md_load", "dlpx_md_run", "dlpx_md_run_ensemble", "dlpx_md_run_ensemble_energies", "dlpx_md_store",
           "dlpx_md_last_output", "dlpx_set_vdw_force_shift", "dlpx_set_vdw_direct", "dlpx_set_exclusions",
           "dlpx_set_per_particle", "dlpx_add_per_particle", "dlpx_set_frozen", "dlpx_compute_lrc_frozen",
           "dlpx_set_timing", "dlpx_get_timing", "dlpx_two_body_forces_noalloc", "dlpx_count_pairs"]


class DlpxError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("dlpx error %d: %s" % (code, msg))
        self.code = code


class Params(C.Structure):
    _fields_ = [("precision", C.c_int32), ("natms_max", C.c_int32), ("ntypes", C.c_int32),
                ("max_list", C.c_int32), ("cutoff", C.c_double), ("padding", C.c_double),
                ("vdw_cutoff", C.c_double), ("electrostatics", C.c_int32), ("nsplines", C.c_int32),
                ("kvec", C.c_int32 * 3), ("alpha", C.c_double), ("epsilon", C.c_double)]


class Atoms(C.Structure):
    _fields_ = [("natms", C.c_int32), ("stride", C.c_int32),
                ("x", C.c_void_p), ("y", C.c_void_p), ("z", C.c_void_p), ("chge", C.c_void_p),
                ("fx", C.c_void_p), ("fy", C.c_void_p), ("fz", C.c_void_p), ("ltype", C.c_void_p)]


class TwoBodyOut(C.Structure):
    _fields_ = [("buffer", C.c_double * 20), ("stress", C.c_double * 9), ("engcpe", C.c_double),
                ("vircpe", C.c_double), ("engsrp", C.c_double), ("virsrp", C.c_double)]


_lib = None


def load():
    """Load libdlpx.so from this package directory; raise if it is not built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            raise ImportError("libdlpx.so is not built: run `python -m dlpx.build` in dl-poly_amd/")
        lib = C.CDLL(LIB)
        lib.dlpx_last_error.restype = C.c_char_p
        lib.dlpx_create.argtypes = [C.POINTER(Params), C.POINTER(C.c_void_p)]
        for name in EXPORTS:
            getattr(lib, name)  # raises AttributeError if the symbol is missing
        _lib = lib
    return _lib


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p)


class Context:
    """One dlpx context = one system layout (atom count, types, cutoffs, SPME mesh)."""

    def __init__(self, sysm, precision=FP64, max_list=0, natms_max=None, tables=None, lrc=None):
        self.lib = load()
        self.sys = sysm
        p = Params(precision=precision, natms_max=natms_max or sysm.natms, ntypes=sysm.ntypes,
                   max_list=max_list, cutoff=sysm.cutoff, padding=sysm.padding,
                   vdw_cutoff=sysm.vdw_cutoff, electrostatics=1 if sysm.electrostatics else 0,
                   nsplines=sysm.nsplines, kvec=(C.c_int32 * 3)(*sysm.kvec),
                   alpha=sysm.alpha if sysm.electrostatics else 1.0, epsilon=sysm.epsilon)
        h = C.c_void_p()
        rc = self.lib.dlpx_create(C.byref(p), C.byref(h))
        if rc != 0:
            raise DlpxError(rc, "dlpx_create failed")
        self.h = h
        self.n = sysm.natms
        self.cell = np.ascontiguousarray(sysm.cell, dtype=np.float64)
        self.q = np.ascontiguousarray(sysm.q)
        self.ltype = np.ascontiguousarray(sysm.types + 1, dtype=np.int32)
        if getattr(sysm, "vdw_force_shift", False):
            self._chk(self.lib.dlpx_set_vdw_force_shift(self.h, C.c_int32(1)))
        if tables is None:
            self._generate_tables()
        else:
            self.set_tables(**tables)
        if lrc is not None:
            self._chk(self.lib.dlpx_set_lrc(self.h, C.c_double(lrc[0]), C.c_double(lrc[1])))
        excl = sysm.exclusions() if hasattr(sysm, "exclusions") else None
        if excl is not None:
            self.set_exclusions(excl)
        frz = getattr(sysm, "frozen", None)
        if frz is not None and np.any(frz):
            f = np.ascontiguousarray(np.asarray(frz, bool).astype(np.int32))
            self._chk(self.lib.dlpx_set_frozen(self.h, C.c_int32(len(f)), _ptr(f)))

    def set_system(self, sysm):
        """Use the context for another configuration of the same species, with
        at most the natms_max it was created for (DL_POLY's natms changes the
        same way between list builds): the host-side arrays switch to sysm and
        the vdW long-range correction is recomputed for its composition and
        volume.  Call link_cell_pairs next."""
        if sysm.ntypes != self.sys.ntypes:
            raise ValueError("set_system needs the same atom types")
        self.sys = sysm
        self.n = sysm.natms
        self.cell = np.ascontiguousarray(sysm.cell, dtype=np.float64)
        self.q = np.ascontiguousarray(sysm.q)
        self.ltype = np.ascontiguousarray(sysm.types + 1, dtype=np.int32)
        num = np.bincount(sysm.types, minlength=sysm.ntypes).astype(np.float64)
        self._chk(self.lib.dlpx_compute_lrc(self.h, _ptr(self.cell), _ptr(num)))

    def set_exclusions(self, excl):
        """Excluded pairs: one sequence of 0-based partner indices per atom
        (symmetric), or None to clear; passed in the reference's
        list_excl(0:mxexcl, 1:natms) layout with 1-based ids."""
        if excl is None:
            self._chk(self.lib.dlpx_set_exclusions(self.h, C.c_int32(0), C.c_int32(0), None))
            return
        mx = max((len(e) for e in excl), default=0)
        arr = np.zeros((len(excl), mx + 1), np.int32)
        for i, e in enumerate(excl):
            arr[i, 0] = len(e)
            arr[i, 1:1 + len(e)] = np.asarray(e, np.int64) + 1
        self._chk(self.lib.dlpx_set_exclusions(self.h, C.c_int32(len(excl)), C.c_int32(mx), _ptr(arr)))

    # -- tables -------------------------------------------------------
    def _generate_tables(self):
        s = self.sys
        nkeys = s.ntypes * (s.ntypes + 1) // 2
        kind = np.full(nkeys, -1, np.int32)
        params = np.zeros((nkeys, VDW_MAX_PARAM))  # vdws%param(1:max_param, key)
        for (ti, tj, k, prm) in s.vdw:
            a, b = ti + 1, tj + 1
            hi, lo = max(a, b), min(a, b)
            key = hi * (hi - 1) // 2 + lo - 1
            kind[key] = k
            params[key, :len(prm)] = prm
        self._chk(self.lib.dlpx_generate_vdw_tables(self.h, C.c_int32(nkeys), _ptr(kind), _ptr(params)))
        if s.electrostatics:
            self._chk(self.lib.dlpx_generate_erfc_tables(self.h))
        num = np.bincount(s.types, minlength=s.ntypes).astype(np.float64)
        frz = getattr(s, "frozen", None)
        if frz is not None and np.any(frz):
            nfz = np.bincount(s.types[np.asarray(frz, bool)], minlength=s.ntypes).astype(np.float64)
            self._chk(self.lib.dlpx_compute_lrc_frozen(self.h, _ptr(self.cell), _ptr(num), _ptr(nfz)))
        else:
            self._chk(self.lib.dlpx_compute_lrc(self.h, _ptr(self.cell), _ptr(num)))
        if getattr(s, "vdw_direct", False):
            self._chk(self.lib.dlpx_set_vdw_direct(self.h, C.c_int32(1)))

    def set_tables(self, max_grid, vdw_pot, vdw_frc, ltp, nsamples=0, erfc=None, erfc_d=None):
        pot = np.ascontiguousarray(vdw_pot, np.float64)
        frc = np.ascontiguousarray(vdw_frc, np.float64)
        ltp = np.ascontiguousarray(ltp, np.int32)
        self._chk(self.lib.dlpx_set_vdw_tables(self.h, C.c_int32(pot.shape[0]), C.c_int32(max_grid),
                                               _ptr(pot), _ptr(frc), _ptr(ltp)))
        if erfc is not None:
            e = np.ascontiguousarray(erfc, np.float64)
            d = np.ascontiguousarray(erfc_d, np.float64)
            self._chk(self.lib.dlpx_set_erfc_tables(self.h, C.c_int32(nsamples), _ptr(e), _ptr(d)))

    def lrc(self):
        e, v = C.c_double(), C.c_double()
        self._chk(self.lib.dlpx_get_lrc(self.h, C.byref(e), C.byref(v)))
        return e.value, v.value

    # -- calls --------------------------------------------------------
    def _chk(self, rc):
        if rc != 0:
            raise DlpxError(rc, self.lib.dlpx_last_error(self.h).decode())

    def _atoms(self, pos, frc=None):
        pos = np.ascontiguousarray(pos, np.float64)
        assert pos.shape == (self.n, 3)
        self._keep = [pos]
        a = Atoms(natms=self.n, stride=24, x=_ptr(pos), y=C.c_void_p(pos.ctypes.data + 8),
                  z=C.c_void_p(pos.ctypes.data + 16), chge=None, ltype=_ptr(self.ltype))
        # charges have their own stride (8) in this binding: pass through a packed copy
        qq = np.zeros((self.n, 3))
        qq[:, 0] = self.q
        self._keep.append(qq)
        a.chge = _ptr(qq)
        if frc is not None:
            assert frc.shape == (self.n, 3) and frc.flags.c_contiguous and frc.dtype == np.float64
            a.fx = _ptr(frc)
            a.fy = C.c_void_p(frc.ctypes.data + 8)
            a.fz = C.c_void_p(frc.ctypes.data + 16)
        return a

    def link_cell_pairs(self, pos):
        a = self._atoms(pos)
        npairs, ibig = C.c_int64(), C.c_int32()
        rc = self.lib.dlpx_link_cell_pairs(self.h, _ptr(self.cell), C.byref(a), C.byref(npairs), C.byref(ibig))
        if rc != 0:
            err = DlpxError(rc, self.lib.dlpx_last_error(self.h).decode())
            err.ibig = ibig.value
            raise err
        return npairs.value, ibig.value

    def get_pairs(self):
        n = C.c_int64()
        self._chk(self.lib.dlpx_get_pairs(self.h, None, C.c_int64(0), C.byref(n)))
        out = np.zeros((max(n.value, 1), 5), np.int32)
        self._chk(self.lib.dlpx_get_pairs(self.h, _ptr(out), C.c_int64(n.value), C.byref(n)))
        return out[:n.value]

    def count_pairs(self, r):
        """Unordered listed pairs closer than r (device count)."""
        n = C.c_int64()
        self._chk(self.lib.dlpx_count_pairs(self.h, C.c_double(r), C.byref(n)))
        return n.value

    def vnl_check(self, pos):
        a = self._atoms(pos)
        u = C.c_int32()
        self._chk(self.lib.dlpx_vnl_check(self.h, C.byref(a), C.byref(u)))
        return bool(u.value)

    def two_body_forces(self, pos, frc=None, per_particle=False, noalloc=False):
        """Forces are added into frc (zeroed here when not given).  With
        per_particle=True the result also holds pp_energy (N) and pp_stress
        (N, 9; element a + 3b of atom i is stats%pp_stress(a+3b+1, i)).  With
        noalloc=True the call goes through dlpx_two_body_forces_noalloc."""
        if frc is None:
            frc = np.zeros((self.n, 3))
        a = self._atoms(pos, frc)
        o = TwoBodyOut()
        self._chk(self.lib.dlpx_set_per_particle(self.h, C.c_int32(1 if per_particle else 0)))
        entry = self.lib.dlpx_two_body_forces_noalloc if noalloc else self.lib.dlpx_two_body_forces
        self._chk(entry(self.h, _ptr(self.cell), C.byref(a), C.byref(o)))
        res = dict(frc=frc, buffer=np.array(o.buffer[:]), stress=np.array(o.stress[:]), engcpe=o.engcpe,
                   vircpe=o.vircpe, engsrp=o.engsrp, virsrp=o.virsrp)
        if per_particle:
            e, st = np.zeros(self.n), np.zeros((self.n, 9))
            self._chk(self.lib.dlpx_add_per_particle(self.h, _ptr(e), _ptr(st)))
            res["pp_energy"], res["pp_stress"] = e, st
        return res

    def md_load(self, pos, vel):
        a = self._atoms(pos)
        v = np.ascontiguousarray(vel, np.float64)
        m = np.zeros((self.n, 3))
        m[:, 0] = self.sys.mass
        self._keep += [v, m]
        self._chk(self.lib.dlpx_md_load(self.h, _ptr(self.cell), C.byref(a), _ptr(v),
                                        C.c_void_p(v.ctypes.data + 8), C.c_void_p(v.ctypes.data + 16), _ptr(m),
                                        C.c_int32(24)))

    def md_run(self, nsteps, tstep, energies=True):
        e = np.zeros((nsteps, 5)) if energies else None
        nreb = C.c_int32()
        self._chk(self.lib.dlpx_md_run(self.h, C.c_int32(nsteps), C.c_double(tstep),
                                       _ptr(e) if energies else None, C.byref(nreb)))
        return e, nreb.value

    def md_store(self):
        out = np.zeros((self.n, 9))
        b = out.ctypes.data
        ptrs = [C.c_void_p(b + 8 * k) for k in range(9)]
        self._chk(self.lib.dlpx_md_store(self.h, *ptrs, C.c_int32(72), C.c_int32(0)))
        return out[:, 0:3].copy(), out[:, 3:6].copy(), out[:, 6:9].copy()

    def md_last_output(self):
        """dlpx_md_last_output: the two-body output of the last step of the last
        md_run(..., energies=True)."""
        o = TwoBodyOut()
        self._chk(self.lib.dlpx_md_last_output(self.h, C.byref(o)))
        return dict(buffer=np.array(o.buffer[:]), stress=np.array(o.stress[:]), engcpe=o.engcpe,
                    vircpe=o.vircpe, engsrp=o.engsrp, virsrp=o.virsrp)

    def set_timing(self, on=True):
        self._chk(self.lib.dlpx_set_timing(self.h, C.c_int32(1 if on else 0)))

    def get_timing(self):
        ms = np.zeros(8)
        self._chk(self.lib.dlpx_get_timing(self.h, _ptr(ms), C.c_int32(8)))
        return ms

    def close(self):
        if getattr(self, "h", None):
            self.lib.dlpx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def md_run_ensemble(contexts, nsteps, tsteps=None, energies=False):
    """dlpx_md_run_ensemble(_energies): nsteps resident NVE steps of every
    context (each loaded with md_load), every stage one launch over all the
    systems.  tsteps defaults to each system's own time step.  Returns the
    list rebuilds per system, and with energies also an array (nsys, nsteps,
    5) of per-step (engcpe, vircpe, engsrp, virsrp, kinetic energy)."""
    n = len(contexts)
    if n == 0:
        return (np.zeros(0, np.int32), np.zeros((0, nsteps, 5))) if energies else np.zeros(0, np.int32)
    ts = np.ascontiguousarray([c.sys.timestep for c in contexts] if tsteps is None else tsteps, np.float64)
    hs = (C.c_void_p * n)(*[c.h for c in contexts])
    nreb = np.zeros(n, np.int32)
    lib = contexts[0].lib
    e = np.zeros((n, nsteps, 5), np.float64)
    rc = lib.dlpx_md_run_ensemble_energies(hs, C.c_int32(n), C.c_int32(nsteps), _ptr(ts),
                                           _ptr(e) if energies else None, _ptr(nreb))
    if rc != 0:
        bad = [c for c in contexts if c.lib.dlpx_last_error(c.h)]
        msg = lib.dlpx_last_error(bad[0].h).decode() if bad else "dlpx_md_run_ensemble failed"
        raise DlpxError(rc, msg)
    return (nreb, e) if energies else nreb
