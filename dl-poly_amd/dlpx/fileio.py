"""DL_POLY file formats used on the force path's boundary.

Writers for CONFIG (``configuration.F90:846`` reader, list-directed reals),
FIELD (``ffield.F90``; ``units internal`` so parameters are taken verbatim) and
new-style CONTROL (``control.F90`` keyword table); readers for REVCON
(``configuration.F90:2601-2615``, ``3g20.10`` records) and the binary REVIVE
(``system.F90:2078-2093``: sequential unformatted records, stpval at record 4,
stress at record 13).
"""
import os
import struct

import numpy as np

from .systems import VDW_NAMES


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


def write_field(path, sysm):
    """One single-atom molecule type per atom type, in type order (the CONFIG
    lists atoms grouped by type in the same order)."""
    counts = np.bincount(sysm.types, minlength=sysm.ntypes)
    order_ok = np.all(np.diff(sysm.types) >= 0)
    if not order_ok:
        raise ValueError("atoms must be grouped by type for the FIELD molecule layout")
    out = [sysm.name, "units internal", "molecules %d" % sysm.ntypes]
    for t, nm in enumerate(sysm.type_names):
        out += [nm, "nummols %d" % counts[t], "atoms 1",
                "%-8s %s %s" % (nm, _num(sysm.masses[t]), _num(sysm.charges[t])), "finish"]
    out.append("vdw %d" % len(sysm.vdw))
    for (ti, tj, key, params) in sysm.vdw:
        out.append("%-8s %-8s %4s %s" % (sysm.type_names[ti], sysm.type_names[tj], VDW_NAMES[key],
                                         " ".join(_num(p) for p in params)))
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
           "vdw_method tabulated",
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
