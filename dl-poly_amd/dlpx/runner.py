"""Run a DL_POLY input directory (CONFIG, FIELD, CONTROL) on the GPU path and
write the files DL_POLY would write for it: STATIS and REVCON.

The run is what DL_POLY's MD driver does for an NVE system of unbonded atoms
with two-body vdW and SPME Coulomb (``drivers.F90:2059-2290``): forces at the
start, then velocity-Verlet steps with the displacement-triggered list
rebuild, resident on the GPU through ``dlpx_md_run``.  Every
``stats_frequency`` steps (and at step 0) the runner writes one STATIS record
with the quantities of ``statistics_collect`` (``statistics.F90:1040-1160``)
that this path produces; the terms of absent interactions (bonds, angles,
shells, constraints, rotational motion) are zero, as in the reference's file.
The reference's own comparator rule (``fileio.statis_equal``) then judges the
file against DL_POLY's.
"""
import os

import numpy as np

from . import fileio, lib


def _cell_angles(cell):
    """alpha, beta, gamma in degrees (dcell, numerics.F90: acos of the cell
    vector cosines, written as stpval(23:25) in statistics.F90:1141-1143)."""
    a, b, c = np.asarray(cell, np.float64).reshape(3, 3)
    cos = lambda u, v: float(np.dot(u, v) / (np.linalg.norm(u) * np.linalg.norm(v)))
    return (np.degrees(np.arccos(cos(b, c))), np.degrees(np.arccos(cos(a, c))), np.degrees(np.arccos(cos(a, b))))


def start_velocities(s):
    """Velocities as DL_POLY starts the run with them (set_temperature,
    temperature.F90:579-620, with the default fixed_com): frozen atoms at rest
    (freeze_atoms), and the centre-of-mass velocity of the other atoms
    (getvom, kinetics.F90:365-419) subtracted from them."""
    v = np.array(s.vel, dtype=np.float64, copy=True)
    m = s.mass
    frz = getattr(s, "frozen", None)
    frz = np.asarray(frz, bool) if frz is not None else np.zeros(s.natms, bool)
    v[frz] = 0.0
    free = ~frz & (m > 1.0e-6)
    tot = float(m[~frz].sum())
    if tot > 0.0:
        vom = (m[~frz, None] * v[~frz]).sum(0) / tot
        v[free] -= vom
    return v


def statis_values(s, out, vel, msd_by_type, engunit=1.0):
    """STATIS columns of one step (statistics.F90:1118-1160, 1337 order):
    stpval(1:27), stpval(0), the pressure tensor (9) and the mean squared
    displacement per atom type.  ``out`` is a two-body output dict (engcpe,
    vircpe, engsrp, virsrp, stress); ``vel`` the velocities of the step."""
    m = s.mass
    frz = getattr(s, "frozen", None)
    nfrz = int(np.count_nonzero(frz)) if frz is not None else 0
    if nfrz:
        vel = np.where(np.asarray(frz, bool)[:, None], 0.0, vel)  # freeze_atoms (configuration.F90:3240-3258)
    engke = 0.5 * float(np.sum(m[:, None] * vel * vel))
    degfre = 3 * s.natms - 3 - 3 * nfrz                         # temperature.F90:110-143 (COM, frozen atoms)
    temp = 2.0 * engke / (fileio.BOLTZ * degfre)
    engcfg = out["engcpe"] + out["engsrp"]
    stpeng = engcfg + engke
    vir = out["vircpe"] + out["virsrp"]
    vol = s.volume
    prs = (2.0 * engke - vir) / (3.0 * vol)                    # statistics.F90:1096-1098
    strkin = np.einsum("i,ia,ib->ab", m, vel, vel).ravel()    # kinetic stress, row-major 3x3
    strtot = np.asarray(out["stress"]) + strkin
    al, be, ga = _cell_angles(s.cell)
    row = np.zeros(28)
    row[1], row[2], row[3] = stpeng / engunit, temp, engcfg / engunit
    row[4], row[5] = out["engsrp"] / engunit, out["engcpe"] / engunit
    row[10] = (stpeng + prs * vol) / engunit                   # enthalpy, fixed cell
    row[12], row[13], row[14] = vir / engunit, out["virsrp"] / engunit, out["vircpe"] / engunit
    row[19] = vol
    row[23], row[24], row[25] = al, be, ga
    row[27] = prs * fileio.PRSUNT
    consv = 0.0                                                 # NVE: no extended-system energy
    return np.concatenate([row[1:28], [consv], strtot * fileio.PRSUNT / vol, msd_by_type])


def run_directory(directory, out_dir=None, precision=lib.FP64, steps=None):
    """Read CONFIG/FIELD/CONTROL from ``directory``, run CONTROL's
    ``time_run`` NVE steps (or ``steps``), and write STATIS and REVCON into
    ``out_dir`` (default: ``directory``).  Returns dict(statis=path,
    revcon=path, rebuilds=count)."""
    s = fileio.system_from_files(directory)
    fld = fileio.read_field(os.path.join(directory, "FIELD"))
    engunit = fileio.ENERGY_UNITS[fld["units"]]
    units_name = {"internal": "DL_POLY Internal UNITS (10 J/mol)", "ev": "electron Volts", "kcal": "kcal/mol",
                  "kj": "kjoule/mol", "k": "Kelvin/Boltzmann"}[fld["units"]]
    nsteps = s.meta["steps"] if steps is None else int(steps)
    every = max(1, s.meta["stats_frequency"])
    out_dir = out_dir or directory
    ntypes = s.ntypes
    counts = np.bincount(s.types, minlength=ntypes).astype(np.float64)

    ctx = lib.Context(s, precision=precision)
    v0 = start_velocities(s)
    try:
        ctx.md_load(s.pos, v0)
        # mean squared displacement as the reference accumulates it for real
        # dynamics (statistics.F90:1162-1176, 1240-1287): xto += v(t_n) dt at
        # every step n >= 1 with the step's full-step velocities, so the run
        # goes one step at a time and copies the velocities back each step
        xto = np.zeros_like(s.pos)
        statis_path = os.path.join(out_dir, "STATIS")
        rebuilds = 0
        with open(statis_path, "w") as f:
            f.write(fileio.statis_header(s.name, units_name))
            out = ctx.md_last_output()
            f.write(fileio.statis_record(0, 0.0, statis_values(s, out, v0, np.zeros(ntypes), engunit)))
            x, v, frc = s.pos, v0, None
            for step in range(1, nsteps + 1):
                _, nreb = ctx.md_run(1, s.timestep, energies=True)
                rebuilds += nreb
                x, v, frc = ctx.md_store()
                xto += v * s.timestep
                if step % every == 0:
                    r2 = np.sum(xto * xto, axis=1)
                    msd = np.bincount(s.types, weights=r2, minlength=ntypes) / np.maximum(counts, 1.0)
                    out = ctx.md_last_output()
                    f.write(fileio.statis_record(step, step * s.timestep, statis_values(s, out, v, msd, engunit)))
        revcon_path = os.path.join(out_dir, "REVCON")
        if frc is None:
            x, v, frc = ctx.md_store()
        fileio.write_revcon(revcon_path, s.name, s.cell, [s.type_names[t] for t in s.types], x, v, frc,
                            nstep=nsteps, tstep=s.timestep, time=nsteps * s.timestep)
    finally:
        ctx.close()
    return dict(statis=statis_path, revcon=revcon_path, rebuilds=rebuilds)


def main(argv=None):
    import argparse
    ap = argparse.ArgumentParser(description="Run a DL_POLY CONFIG/FIELD/CONTROL directory on the GPU path "
                                             "and write STATIS and REVCON.")
    ap.add_argument("directory")
    ap.add_argument("--out", default=None, help="output directory (default: the input directory)")
    ap.add_argument("--mixed", action="store_true", help="FP32 arithmetic with FP64 accumulation")
    ap.add_argument("--steps", type=int, default=None, help="override CONTROL's time_run")
    a = ap.parse_args(argv)
    r = run_directory(a.directory, out_dir=a.out, precision=lib.MIXED if a.mixed else lib.FP64, steps=a.steps)
    print("wrote %s and %s (%d list rebuilds)" % (r["statis"], r["revcon"], r["rebuilds"]))


if __name__ == "__main__":
    main()
