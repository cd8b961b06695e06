"""DL_POLY file formats used on the force path's boundary.

Readers and writers of the files a DL_POLY run of this path reads and
writes, so runs of this library can be set up and judged with DL_POLY's own
inputs and comparators:

* CONFIG / REVCON: ``read_config`` (``configuration.F90:846`` reader:
  title, ``levcfg imcon natms`` record, cell for imcon > 0, then per atom a
  name/index record and position, velocity, force records by levcfg) and
  ``write_config`` / ``write_revcon`` (``configuration.F90:2540-2615``:
  ``(4i10,1p,2e16.7)`` header, ``(a8,i10)`` and ``3g20.10`` records);
* FIELD: ``read_field`` (``ffield.F90:150-5365``, the subset this path
  supports: energy units, molecule types of unbonded atoms, the ``12-6``,
  ``lj``, ``nm``, ``buck``, ``bhm``, ``hbnd``, ``mors``, ``wca``, ``dpd``,
  ``ljc``, ``mstw``, ``ryd``, ``sand``, ``ndpd``, ``snm`` and ``14-7`` vdW forms) and
  ``write_field``;
* CONTROL: ``read_control`` / ``write_control`` (new-style keyword file,
  ``control.F90``);
* STATIS: ``statis_header`` / ``statis_record`` (``statistics.F90:1020-1027,
  1333-1347``: ``(i10,1p,e14.6,0p,i10,/,(1p,5e14.6))`` records), ``read_statis``
  and ``statis_equal``, the relative-precision rule of the reference's test
  harness (``utils/beetest:39-48``);
* REVIVE (binary restart): ``read_revive`` (``system.F90:2078-2093``:
  sequential unformatted records, stpval at record 4, stress at record 13).
"""
import os
import struct

import numpy as np

from .systems import VDW_LJF, VDW_NAMES, VDW_NPARAMS, VDW_WCA, ljf_field_words, ljf_params, wca_params


# constants.F90:104,112-114,123: Boltzmann constant and energy units in
# internal units (10 J/mol), pressure unit (internal -> katm)
BOLTZ = 8.31451115e-1
ENERGY_UNITS = {"internal": 1.0, "ev": 9648.530821, "kcal": 418.4, "kj": 100.0, "k": BOLTZ}
PRSUNT = 1.63882576e-1


def _num(x):
    # shortest string that round-trips to the same double
    return repr(float(x))


def write_config(path, sysm, levcfg=1, title=None):
    n = sysm.natms
    imcon = 1 if _cubic(sysm.cell) else (2 if _orthorhombic(sysm.cell) else 3)
    lines = [(title or sysm.name)[:72], "%10d%10d%10d" % (levcfg, imcon, n)]
    c = sysm.cell.reshape(3, 3)
    for r in range(3):
        lines.append(" ".join(_num(v) for v in c[r]))
    for i in range(n):
        lines.append("%-8s%10d" % (sysm.type_names[sysm.types[i]], i + 1))
        lines.append(" ".join(_num(v) for v in sysm.pos[i]))
        if levcfg > 0:
            lines.append(" ".join(_num(v) for v in sysm.vel[i]))
    for ln in lines:
        if len(ln) > 72:
            raise ValueError("CONFIG record longer than 72 characters: %r" % ln)
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")


def _orthorhombic(cell):
    c = cell.reshape(3, 3)
    return bool(np.all(c == np.diag(np.diag(c))))


def _cubic(cell):
    c = cell.reshape(3, 3)
    return np.allclose(c, np.diag(np.diag(c))) and c[0, 0] == c[1, 1] == c[2, 2]


def _site_frozen(sysm, start, nummols, nsites, name):
    """Frozen flags of the sites of a molecule type whose nummols copies start
    at atom ``start`` (a FIELD gives one flag per site, shared by all copies)."""
    frz = getattr(sysm, "frozen", None)
    if frz is None:
        return [False] * nsites
    f = np.asarray(frz, bool)[start:start + nummols * nsites].reshape(max(nummols, 1), nsites)
    if not np.all(f == f[0]):
        raise ValueError("molecule type %r: copies differ in which sites are frozen" % name)
    return [bool(x) for x in f[0]]


def write_field(path, sysm):
    """Without ``sysm.molecules``: one single-atom molecule type per atom type,
    in type order (the CONFIG lists atoms grouped by type in the same order).
    With them: each molecule type with its sites, and its bonds and angles as
    zero-strength harmonic terms (``harm``), which exclude their pairs from the
    two-body sum and add no force."""
    mols = getattr(sysm, "molecules", None)
    if mols:
        out = [sysm.name, "units internal", "molecules %d" % len(mols)]
        base = 0
        for m in mols:
            got = sysm.types[base:base + m.nummols * len(m.sites)]
            if not np.array_equal(got, np.tile(np.asarray(m.sites, np.int32), m.nummols)):
                raise ValueError("atoms do not follow molecule type %r's site order" % m.name)
            base += m.nummols * len(m.sites)
            out += [m.name, "nummols %d" % m.nummols, "atoms %d" % len(m.sites)]
            fz = _site_frozen(sysm, base - m.nummols * len(m.sites), m.nummols, len(m.sites), m.name)
            for t, f in zip(m.sites, fz):
                out.append("%-8s %s %s" % (sysm.type_names[t], _num(sysm.masses[t]), _num(sysm.charges[t])) +
                           (" 1 1" if f else ""))
            if m.bonds:
                out.append("bonds %d" % len(m.bonds))
                out += ["harm %d %d 0.0 1.0" % (a + 1, b + 1) for (a, b) in m.bonds]
            if m.angles:
                out.append("angles %d" % len(m.angles))
                out += ["harm %d %d %d 0.0 109.47" % (a + 1, b + 1, c + 1) for (a, b, c) in m.angles]
            out.append("finish")
        if base != sysm.natms:
            raise ValueError("molecules describe %d atoms, the system has %d" % (base, sysm.natms))
    else:
        counts = np.bincount(sysm.types, minlength=sysm.ntypes)
        order_ok = np.all(np.diff(sysm.types) >= 0)
        if not order_ok:
            raise ValueError("atoms must be grouped by type for the FIELD molecule layout")
        out = [sysm.name, "units internal", "molecules %d" % sysm.ntypes]
        base = 0
        for t, nm in enumerate(sysm.type_names):
            fz = _site_frozen(sysm, base, int(counts[t]), 1, nm)
            base += int(counts[t])
            out += [nm, "nummols %d" % counts[t], "atoms 1",
                    "%-8s %s %s" % (nm, _num(sysm.masses[t]), _num(sysm.charges[t])) + (" 1 1" if fz[0] else ""),
                    "finish"]
    out.append("vdw %d" % len(sysm.vdw))
    for (ti, tj, key, params) in sysm.vdw:
        out.append("%-8s %-8s %4s %s" % (sysm.type_names[ti], sysm.type_names[tj], VDW_NAMES[key],
                                         " ".join(_num(p) for p in (ljf_field_words(params) if key == VDW_LJF
                                                                     else params[:_VDW_FIELD_NPARAMS[key]]))))
    out.append("close")
    with open(path, "w") as f:
        f.write("\n".join(out) + "\n")


def write_control(path, sysm, steps=0, print_every=1, extra=()):
    out = ["title %s" % sysm.name,
           "simulation_method md",
           "ensemble nve",
           "temperature %s K" % _num(max(sysm.temperature, 1.0)),
           "timestep %s ps" % _num(sysm.timestep),
           "time_run %d steps" % steps,
           "time_equilibration 0 steps",
           "restart noscale",
           "cutoff %s ang" % _num(sysm.cutoff),
           "padding %s ang" % _num(sysm.padding),
           "vdw_cutoff %s ang" % _num(sysm.vdw_cutoff),
           "vdw_method %s" % ("direct" if getattr(sysm, "vdw_direct", False) else "tabulated"),
           "print_frequency %d steps" % print_every,
           "stats_frequency %d steps" % print_every,
           "time_job 100000 s",
           "time_close 100 s"]
    if sysm.electrostatics:
        out += ["coul_method spme",
                "coul_dielectric_constant %s" % _num(sysm.epsilon),
                "spme_alpha %s ang^-1" % _num(sysm.alpha),
                "spme_kvec [%d %d %d]" % tuple(sysm.kvec),
                "spme_nsplines %d" % sysm.nsplines]
    else:
        out.append("coul_method off")
    if getattr(sysm, "vdw_force_shift", False):
        out.append("vdw_force_shift on")
    out += list(extra)
    with open(path, "w") as f:
        f.write("\n".join(out) + "\n")


def write_inputs(directory, sysm, steps=0, print_every=1, extra=()):
    os.makedirs(directory, exist_ok=True)
    write_config(os.path.join(directory, "CONFIG"), sysm, levcfg=1)
    write_field(os.path.join(directory, "FIELD"), sysm)
    write_control(os.path.join(directory, "CONTROL"), sysm, steps=steps,
                  print_every=print_every, extra=extra)


def read_revcon(path):
    """Returns dict(cell (9,), pos, vel, frc (N,3) in CONFIG index order)."""
    with open(path) as f:
        lines = f.read().splitlines()
    levcfg, imcon, n = (int(v) for v in lines[1].split()[:3])
    cell = np.array([float(v) for ln in lines[2:5] for v in ln.split()])
    per = 2 + (levcfg >= 1) + (levcfg >= 2)
    pos = np.zeros((n, 3)); vel = np.zeros((n, 3)); frc = np.zeros((n, 3))
    k = 5
    for _ in range(n):
        idx = int(lines[k].split()[1]) - 1
        pos[idx] = [float(v) for v in lines[k + 1].split()]
        if levcfg >= 1:
            vel[idx] = [float(v) for v in lines[k + 2].split()]
        if levcfg >= 2:
            frc[idx] = [float(v) for v in lines[k + 3].split()]
        k += per
    return dict(cell=cell, pos=pos, vel=vel, frc=frc, levcfg=levcfg)


def read_fortran_records(path):
    """Sequential unformatted file: [int32 len][payload][int32 len] per record."""
    recs = []
    with open(path, "rb") as f:
        data = f.read()
    off = 0
    while off < len(data):
        (ln,) = struct.unpack_from("<i", data, off)
        payload = data[off + 4: off + 4 + ln]
        (ln2,) = struct.unpack_from("<i", data, off + 4 + ln)
        if ln2 != ln:
            raise ValueError("corrupt record marker at %d" % off)
        recs.append(payload)
        off += 8 + ln
    return recs


# statistics.F90:1118-1145 slots of stats%stpval used here
STP_ENGCFG, STP_ENGSRP, STP_ENGCPE, STP_VIRSRP, STP_VIRCPE = 3, 4, 5, 13, 14


def read_revive(path):
    recs = read_fortran_records(path)
    f64 = lambda r: np.frombuffer(r, dtype="<f8")
    stpval = f64(recs[3])
    stress = f64(recs[12])
    head = f64(recs[1])
    return dict(stpval=stpval, stress=stress.copy(), nstep=int(head[0]),
                engsrp=stpval[STP_ENGSRP], engcpe=stpval[STP_ENGCPE],
                virsrp=stpval[STP_VIRSRP], vircpe=stpval[STP_VIRCPE],
                engcfg=stpval[STP_ENGCFG])


# ---------------------------------------------------------------- CONFIG --

def read_config(path):
    """CONFIG or REVCON (configuration.F90:846-1100 reader).  Returns dict(title,
    levcfg, imcon, cell (9,), names (N,), index (N,), pos, vel, frc (N, 3))
    in file order; vel/frc are zero when levcfg does not provide them."""
    with open(path) as f:
        lines = f.read().splitlines()
    title = lines[0].rstrip()
    head = lines[1].split()
    levcfg, imcon = int(head[0]), int(head[1])
    n_decl = int(head[2]) if len(head) > 2 else None
    k = 2
    cell = np.zeros(9)
    if imcon > 0:
        cell = np.array([float(v) for ln in lines[2:5] for v in ln.split()[:3]])
        k = 5
    per = 2 + (levcfg >= 1) + (levcfg >= 2)
    names, index, pos, vel, frc = [], [], [], [], []
    while k + per - 1 < len(lines) and lines[k].strip():
        rec = lines[k].split()
        names.append(rec[0])
        index.append(int(rec[1]) if len(rec) > 1 else len(names))
        pos.append([float(v) for v in lines[k + 1].split()[:3]])
        vel.append([float(v) for v in lines[k + 2].split()[:3]] if levcfg >= 1 else [0.0] * 3)
        frc.append([float(v) for v in lines[k + 3].split()[:3]] if levcfg >= 2 else [0.0] * 3)
        k += per
    if n_decl is not None and n_decl != len(names):
        raise ValueError("%s declares %d atoms, holds %d" % (path, n_decl, len(names)))
    return dict(title=title, levcfg=levcfg, imcon=imcon, cell=cell, names=names,
                index=np.array(index, np.int64), pos=np.array(pos), vel=np.array(vel), frc=np.array(frc))


def _g20(x):
    # a 20-character field with 10 significant digits (the reference writes 3g20.10)
    return "%20.10G" % float(x)


def write_revcon(path, title, cell, names, pos, vel=None, frc=None, nstep=0, tstep=0.0, time=0.0, imcon=None):
    """REVCON in the reference's layout (configuration.F90:2540-2615): levcfg 2
    when forces are given, 1 with velocities only, 0 otherwise."""
    n = len(names)
    levcfg = 2 if frc is not None else (1 if vel is not None else 0)
    if imcon is None:
        imcon = 1 if _cubic(np.asarray(cell)) else (2 if _orthorhombic(np.asarray(cell)) else 3)
    out = ["%-72s" % title[:72], "%10d%10d%10d%10d%16.7E%16.7E" % (levcfg, imcon, n, nstep, tstep, time)]
    c = np.asarray(cell, np.float64).reshape(3, 3)
    for r in range(3):
        out.append("%20.10f%20.10f%20.10f" % tuple(c[r]))
    for i in range(n):
        out.append("%-8s%10d" % (names[i], i + 1))
        out.append("".join(_g20(v) for v in pos[i]))
        if levcfg >= 1:
            out.append("".join(_g20(v) for v in vel[i]))
        if levcfg >= 2:
            out.append("".join(_g20(v) for v in frc[i]))
    with open(path, "w") as f:
        f.write("\n".join(out) + "\n")


# ----------------------------------------------------------------- FIELD --

_VDW_KEYS = {"12-6": 1, "lj": 2, "nm": 3, "buck": 4, "bhm": 5, "hbnd": 6, "mors": 8, "wca": 9, "dpd": 10,
             "ljc": 12, "mstw": 13, "ryd": 14, "sand": 22, "ndpd": 23, "snm": 7, "14-7": 11, "zbl": 15,
             "zbls": 16, "zblb": 17, "mlj": 18, "mbuc": 19, "m126": 20, "ljf": 21, "sw": 24}
# parameters that carry energy units, per vdW form (ffield.F90:3757-3834:
# parpot(1) except for the ZBL forms, whose parpot(1) is a charge; 12-6,
# hbnd, ryd and m126 parpot(2); buck and mbuc parpot(3); bhm parpot(4:5);
# mstw parpot(4); zbls parpot(5); zblb parpot(5) and parpot(7)); LJF converts
# its own (ljf_params)
_VDW_ENERGY_PARAMS = {1: (0, 1), 2: (0,), 3: (0,), 4: (0, 2), 5: (0, 3, 4), 6: (0, 1), 8: (0,), 9: (0,), 10: (0,),
                      12: (0,), 13: (0, 3), 14: (0, 1), 22: (0,), 23: (0,), 7: (0,), 11: (0,), 15: (), 16: (4,),
                      17: (4, 6), 18: (0,), 19: (0, 2), 20: (0, 1), 21: (), 24: (0,)}
_ZERO_PLUS = 2.2250738585072014e-308  # Tiny(1.0_wp), the reference's zero_plus
# parameters a FIELD record gives (WCA's cut is derived, ffield.F90:3788)
_VDW_FIELD_NPARAMS = {**VDW_NPARAMS, VDW_WCA: 3}


def _bonded_block(lines, i, count, natoms, kind, mname):
    """``count`` bond (natoms 2) or angle (natoms 3) records from lines[i]:
    ``key i j [k] params``.  This path computes no bonded forces, so only
    zero-strength harmonic terms are accepted: ``harm`` (excludes its pairs,
    ffield.F90:1348-1357, 1592-1601) and ``-hrm`` (restrained: excludes
    nothing).  Returns the site tuples (0-based) of the excluding terms."""
    out = []
    for _ in range(count):
        w = lines[i].split()
        i += 1
        key = w[0].lower()[:4]
        idx = [int(v) - 1 for v in w[1:1 + natoms]]
        prm = [float(v) for v in w[1 + natoms:]]
        if key not in ("harm", "-hrm"):
            raise ValueError("%s %r in molecule %r: bonded forces are not part of this path" % (kind, w[0], mname))
        if not prm or prm[0] != 0.0:
            raise ValueError("%s in molecule %r has a nonzero force constant: bonded forces are not part of "
                             "this path (only zero-strength terms, which just exclude pairs)" % (kind, mname))
        if min(idx) < 0:
            raise ValueError("%s in molecule %r names site %d" % (kind, mname, min(idx) + 1))
        if key == "harm":
            out.append(tuple(idx))
    return out, i


def read_field(path):
    """FIELD (ffield.F90:150-5365), the subset of this path: energy units;
    molecule types of atoms (``name mass charge [repeat [frozen]]``; frozen
    atoms, config%lfrzn, are kept in place and pairs of them leave the list) with
    zero-strength harmonic ``bonds`` and ``angles`` that only exclude pairs;
    the vdW forms the library generates (``12-6``, ``lj``, ``nm``, ``buck``,
    ``bhm``, ``hbnd``, ``mors``).  Bonded forces, constraints, shells, other
    vdW forms and many-body terms raise ValueError.
    Returns dict(title, units, molecules [(name, nummols, [(atom, mass, charge)])],
    topology [(bonds, angles) per molecule type, 0-based sites], frozen [per
    molecule type, a flag per site],
    vdw [(atom1, atom2, key, params in internal units)])."""
    with open(path) as f:
        raw = [ln.split("#")[0].rstrip() for ln in f.read().splitlines()]
    title = raw[0].strip()
    lines = [ln for ln in raw[1:] if ln.strip()]
    units = "internal"
    unit = 1.0
    molecules, vdw, topology, frozen = [], [], [], []
    i = 0

    def words(ln):
        return ln.split()

    while i < len(lines):
        w = words(lines[i])
        key = w[0].lower()
        if key == "units":
            u = w[1].lower() if len(w) > 1 else "internal"
            for name, val in (("ev", ENERGY_UNITS["ev"]), ("kcal", ENERGY_UNITS["kcal"]), ("kj", ENERGY_UNITS["kj"]),
                              ("internal", 1.0)):
                if u.startswith(name):
                    units, unit = name, val
                    break
            else:
                if u.startswith("k"):
                    units, unit = "k", ENERGY_UNITS["k"]
                else:
                    raise ValueError("unknown energy units %r" % u)
            i += 1
        elif key.startswith("molecul"):
            nmol = int(w[-1])
            i += 1
            for _ in range(nmol):
                mname = lines[i].strip()
                i += 1
                nummols, atoms, bonds, angles, frz = None, [], [], [], []
                while True:
                    w = words(lines[i])
                    k2 = w[0].lower()
                    if k2.startswith("nummol"):
                        nummols = int(w[-1])
                        i += 1
                    elif k2 == "atoms":
                        na = int(w[-1])
                        i += 1
                        got = 0
                        while got < na:
                            a = words(lines[i])
                            rep = int(a[3]) if len(a) > 3 else 1
                            fflag = int(a[4]) if len(a) > 4 else 0
                            for _r in range(max(rep, 1)):
                                atoms.append((a[0], float(a[1]), float(a[2])))
                                frz.append(fflag != 0)
                            got += max(rep, 1)
                            i += 1
                    elif k2 == "bonds":
                        bonds, i = _bonded_block(lines, i + 1, int(w[-1]), 2, "bond", mname)
                    elif k2 == "angles":
                        angles, i = _bonded_block(lines, i + 1, int(w[-1]), 3, "angle", mname)
                    elif k2 == "finish":
                        i += 1
                        break
                    else:
                        raise ValueError("FIELD section %r is not supported by this path" % w[0])
                if nummols is None:
                    raise ValueError("molecule %r has no nummols record" % mname)
                for t in bonds + angles:
                    if max(t) >= len(atoms):
                        raise ValueError("a bonded term of molecule %r names site %d of %d" % (mname, max(t) + 1,
                                                                                             len(atoms)))
                molecules.append((mname, nummols, atoms))
                topology.append((bonds, angles))
                frozen.append(frz)
        elif key == "vdw":
            nv = int(w[-1])
            i += 1
            for _ in range(nv):
                a = words(lines[i])
                form = a[2].lower()
                if form not in _VDW_KEYS:
                    raise ValueError("vdW form %r is not supported by this path" % a[2])
                kind = _VDW_KEYS[form]
                params = [float(v) for v in a[3:]]
                # absent words read as 0; words past the form's own parameters
                # are read by the reference (vdws%max_param of them) but unused
                params += [0.0] * (_VDW_FIELD_NPARAMS[kind] - len(params))
                params = params[:_VDW_FIELD_NPARAMS[kind]]
                if kind == 4 and abs(params[1]) <= _ZERO_PLUS:
                    # ffield.F90:3745-3751: a zero rho is 1 when A is zero too
                    # (pure dispersion), else error 467
                    if abs(params[0]) <= _ZERO_PLUS:
                        params[1] = 1.0
                    else:
                        raise ValueError("buck vdW pair %s-%s: rho is zero with A nonzero (DL_POLY error 467)"
                                         % (a[0], a[1]))
                if kind in (15, 16, 17):
                    # parpot(1) is scaled to internal units and back (ffield.F90:3757,3806-3816)
                    params[0] = params[0] * unit / unit
                for j in _VDW_ENERGY_PARAMS[kind]:
                    params[j] *= unit
                if kind == VDW_LJF:
                    params = list(ljf_params(params[0] * unit, params[1], params[2]))
                if kind == VDW_WCA:
                    params = list(wca_params(*params[:3]))
                if kind == 7 and params[1] <= params[2]:  # ffield.F90:3753-3755, errors 470 and 468
                    raise ValueError("snm vdW pair %s-%s: n must exceed m" % (a[0], a[1]))
                if kind == 7 and params[4] < params[3]:
                    raise ValueError("snm vdW pair %s-%s: r_trunc is below r0" % (a[0], a[1]))
                vdw.append((a[0], a[1], kind, tuple(params)))
                i += 1
        elif key == "close":
            break
        else:
            raise ValueError("FIELD directive %r is not supported by this path" % w[0])
    return dict(title=title, units=units, molecules=molecules, topology=topology, frozen=frozen, vdw=vdw)


# --------------------------------------------------------------- CONTROL --

_TIME_UNITS = {"ps": 1.0, "fs": 1.0e-3, "ns": 1.0e3}
_LENGTH_UNITS = {"ang": 1.0, "nm": 10.0, "bohr": 0.52917721067}


def read_control(path):
    """New-style CONTROL (control.F90 keyword table), the keys this path uses:
    title, ensemble (nve), timestep, time_run, time_equilibration,
    stats_frequency, cutoff, padding, vdw_cutoff, vdw_method (tabulated /
    direct), vdw_force_shift,
    coul_method (spme / off), coul_dielectric_constant, spme_alpha, spme_kvec,
    spme_nsplines, temperature.  Other keys are kept verbatim in ``other``."""
    out = dict(title="", ensemble="nve", timestep=None, steps=0, equilibration=0, stats_frequency=1,
               cutoff=None, padding=0.0, vdw_cutoff=None, coul_method="off", epsilon=1.0, alpha=None,
               kvec=None, nsplines=8, temperature=0.0, vdw_force_shift=False, vdw_method="tabulated", other={})
    with open(path) as f:
        for ln in f:
            ln = ln.split("#")[0].strip()
            if not ln:
                continue
            w = ln.split()
            key = w[0].lower()
            vals = w[1:]

            def length(v):
                return float(v[0]) * _LENGTH_UNITS.get(v[1].lower(), 1.0) if len(v) > 1 else float(v[0])

            if key == "title":
                out["title"] = ln[len(w[0]):].strip()
            elif key == "ensemble":
                out["ensemble"] = vals[0].lower()
            elif key == "timestep":
                out["timestep"] = float(vals[0]) * (_TIME_UNITS.get(vals[1].lower(), 1.0) if len(vals) > 1 else 1.0)
            elif key == "time_run":
                out["steps"] = int(float(vals[0]))
            elif key == "time_equilibration":
                out["equilibration"] = int(float(vals[0]))
            elif key == "stats_frequency":
                out["stats_frequency"] = int(float(vals[0]))
            elif key == "cutoff":
                out["cutoff"] = length(vals)
            elif key == "padding":
                out["padding"] = length(vals)
            elif key == "vdw_cutoff":
                out["vdw_cutoff"] = length(vals)
            elif key == "coul_method":
                out["coul_method"] = vals[0].lower()
            elif key == "coul_dielectric_constant":
                out["epsilon"] = float(vals[0])
            elif key == "spme_alpha":
                out["alpha"] = float(vals[0])
            elif key == "spme_kvec":
                nums = ln[ln.index(w[1]):].replace("[", " ").replace("]", " ").split()
                out["kvec"] = tuple(int(float(v)) for v in nums[:3])
            elif key == "spme_nsplines":
                out["nsplines"] = int(float(vals[0]))
            elif key == "temperature":
                out["temperature"] = float(vals[0])
            elif key == "vdw_method":
                out["vdw_method"] = vals[0].lower()
            elif key == "vdw_force_shift":
                out["vdw_force_shift"] = (vals[0].lower() if vals else "on") in ("on", "true", "yes", "1")
            else:
                out["other"][key] = " ".join(vals)
    if out["vdw_cutoff"] is None:
        out["vdw_cutoff"] = out["cutoff"]
    return out


def system_from_files(directory):
    """A ``System`` from DL_POLY's CONFIG, FIELD and CONTROL in a directory.
    Atom types are the FIELD's unique atom names in order of first appearance
    (sites%unique_atom); the CONFIG must list the atoms in FIELD order
    (molecule type by molecule type, nummols copies each)."""
    from .systems import System
    cfg = read_config(os.path.join(directory, "CONFIG"))
    fld = read_field(os.path.join(directory, "FIELD"))
    ctl = read_control(os.path.join(directory, "CONTROL"))
    if ctl["vdw_method"] not in ("tabulated", "direct"):
        raise ValueError("vdw_method %r: this path has tabulated or direct vdW" % ctl["vdw_method"])
    if ctl["ensemble"] != "nve":
        raise ValueError("ensemble %r: this path integrates NVE only" % ctl["ensemble"])
    if cfg["imcon"] not in (1, 2):
        raise ValueError("imcon %d: this path supports cubic and orthorhombic cells" % cfg["imcon"])
    type_names, masses, charges, site_types = [], [], [], []
    for (_m, nummols, atoms) in fld["molecules"]:
        for _k in range(nummols):
            for (nm, mass, q) in atoms:
                if nm not in type_names:
                    type_names.append(nm)
                    masses.append(mass)
                    charges.append(q)
                t = type_names.index(nm)
                if masses[t] != mass or charges[t] != q:
                    raise ValueError("atom %r appears with different mass or charge" % nm)
                site_types.append(t)
    n = len(cfg["names"])
    if len(site_types) != n:
        raise ValueError("FIELD describes %d atoms, CONFIG holds %d" % (len(site_types), n))
    for i, nm in enumerate(cfg["names"]):
        if type_names[site_types[i]] != nm:
            raise ValueError("CONFIG atom %d is %r, FIELD expects %r" % (i + 1, nm, type_names[site_types[i]]))
    vdw = []
    for (a1, a2, kind, params) in fld["vdw"]:
        if a1 not in type_names or a2 not in type_names:
            raise ValueError("vdW pair %s-%s names an unknown atom" % (a1, a2))
        vdw.append((type_names.index(a1), type_names.index(a2), kind, params))
    spme = ctl["coul_method"].startswith("spme")
    if not spme and not ctl["coul_method"].startswith("off"):
        raise ValueError("coul_method %r: this path has SPME or no electrostatics" % ctl["coul_method"])
    if spme and (ctl["alpha"] is None or ctl["kvec"] is None):
        raise ValueError("SPME needs spme_alpha and spme_kvec in CONTROL")
    kw = {}
    if spme:
        kw = dict(alpha=ctl["alpha"], kvec=ctl["kvec"], nsplines=ctl["nsplines"])
    site_frozen = []
    for (_m, nummols, atoms), fz in zip(fld["molecules"], fld["frozen"]):
        site_frozen += list(fz) * nummols
    if any(site_frozen):
        kw["frozen"] = np.array(site_frozen, bool)
    if any(b or a for (b, a) in fld["topology"]):
        from .systems import Molecule
        kw["molecules"] = [Molecule(name=m, nummols=nm, sites=[type_names.index(x[0]) for x in atoms],
                                    bonds=list(b), angles=list(a))
                           for (m, nm, atoms), (b, a) in zip(fld["molecules"], fld["topology"])]
    return System(name=cfg["title"].strip() or fld["title"], cell=cfg["cell"].copy(), pos=cfg["pos"].copy(),
                  vel=cfg["vel"].copy(), types=np.array(site_types, np.int32), type_names=type_names,
                  masses=np.array(masses), charges=np.array(charges), vdw=vdw, cutoff=ctl["cutoff"],
                  padding=ctl["padding"], vdw_cutoff=ctl["vdw_cutoff"], electrostatics=spme,
                  epsilon=ctl["epsilon"], timestep=ctl["timestep"], temperature=ctl["temperature"],
                  vdw_force_shift=ctl["vdw_force_shift"],
                  vdw_direct=ctl["vdw_method"] == "direct",
                  meta=dict(steps=ctl["steps"], stats_frequency=ctl["stats_frequency"], levcfg=cfg["levcfg"],
                            imcon=cfg["imcon"]), **kw)


# ---------------------------------------------------------------- STATIS --

def statis_header(title, units="DL_POLY Internal UNITS (10 J/mol)"):
    """The two header records (statistics.F90:1005-1026)."""
    return "title: '%-72s'\nenergy unitS: %s\n" % (title[:72], units)


def statis_record(nstep, time, values):
    """One record: (i10,1p,e14.6,0p,i10,/,(1p,5e14.6)) of the values in
    column order (statistics.F90:1337)."""
    v = list(values)
    out = ["%10d%14.6E%10d" % (nstep, time, len(v))]
    for k in range(0, len(v), 5):
        out.append("".join("%14.6E" % x for x in v[k:k + 5]))
    return "\n".join(out) + "\n"


def read_statis(path):
    """Records of a STATIS file: dict nstep -> (time, values array)."""
    with open(path) as f:
        lines = f.read().splitlines()
    recs = {}
    k = 2
    while k < len(lines):
        h = lines[k].split()
        if not h:
            k += 1
            continue
        nstep, time, num = int(h[0]), float(h[1]), int(h[2])
        vals = []
        k += 1
        while len(vals) < num:
            vals += [float(x) for x in lines[k].split()]
            k += 1
        recs[nstep] = (time, np.array(vals[:num]))
    return recs


def statis_equal(actual, expected, precision):
    """utils/beetest:39-48 EqualToRelativePrecision: both zero, or an absolute
    test when the expected value is zero, else |a - e| / |e| < precision."""
    if actual == 0.0 and expected == 0.0:
        return True
    if expected == 0.0:
        return abs(actual - expected) < abs(precision)
    return abs((actual - expected) / expected) < abs(precision)


# STATIS columns (1-based, as the reference's utils/Scripts read them)
STATIS_COLUMNS = {"eng_tot": 1, "temp_tot": 2, "eng_cfg": 3, "eng_src": 4, "eng_cou": 5, "eng_pv": 10,
                  "vir_cfg": 12, "vir_src": 13, "vir_cou": 14, "volume": 19, "press": 27}
