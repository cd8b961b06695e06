#!/usr/bin/env python3
"""Throughput of the per-step force path on one MI355X.

Workload (default): molten NaCl, 64,000 ions (20^3 rocksalt cells, lattice
5.64 A expanded 8%, 1200 K, jitter U(-0.1, 0.1) A), Born-Huggins-Meyer
(Tosi-Fumi) + SPME (rc 12 A, alpha 0.28 A^-1, mesh 128^3, B-spline order 8),
Verlet padding 0.5 A, velocity-Verlet NVE with dt = 2 fs.  A step is one
velocity-Verlet step: neighbour-list check (and rebuild when triggered), the
full two-body + SPME force evaluation, and both integrator half-steps, all
resident in HBM.

Prints one JSON line: value = mixed-precision atom-steps/s; the FP64 path's
rate, the mixed/FP64 ratio and the mixed path's force error against FP64 are
reported beside it, with a per-stage roofline and the compiled reference
(DL_POLY 5.1, serial) timed on the host as the CPU baseline.
"""
import argparse
import json
import math
import os
import re
import subprocess
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "dl-poly_amd"))

import numpy as np  # noqa: E402

from dlpx import fileio, lib, systems  # noqa: E402

METRIC = "atom-steps/s (pair+SPME), mixed vs FP64 on 1 MI355X at RMS force err≤1e-5"
HBM_PEAK = 8.0e12            # B/s
VALU_PEAK = {lib.FP64: 78.6e12, lib.MIXED: 157.3e12}  # FLOP/s, vector FP64 / FP32


# kernels of each stage (mixed precision, spline order 8, 128^3 mesh), as the
# stage-timed pass launches them (single-system kernels; resident runs launch
# the _batch forms of the same bodies)
KERNELS = {"pair": "k_pair_comb<8, 1024, true>", "spread": "k_spread_cells<float, 8>",
           "fft": "k_plane/k_lines_zconv/k_plane<float, 16, 8>",
           "gather": "k_gather<float, 8>"}
# the FFT stage's kernels and their launches per stage: x and y passes fused
# per z-plane (forward, inverse) around the z pass with the influence function
FFT_KERNELS = [("k_plane<float, 16, 8, false>", 1), ("k_lines_zconv<float, 16, 8, 16>", 1),
               ("k_plane<float, 16, 8, true>", 1)]
BATCH_PAIR_KERNEL = "k_pair_comb_batch<8, 1024, true>"
PAIR_KERNEL_VDW = "k_pair<float, 8, 512, false, false, false, false>"  # mixed pair kernel of a vdW-only system


def traffic_from_profile(workload, kernel):
    """HBM bytes per launch of a kernel from the committed rocprofv3 --pmc
    summary (FETCH_SIZE + WRITE_SIZE per dispatch, each scaled by the same
    counter's reading for a plain copy kernel of known size), or None."""
    path = os.path.join(ROOT, "profiles", "traffic_%s.txt" % workload)
    if not os.path.exists(path):
        return None
    for ln in open(path):
        if ln.startswith("#"):
            continue
        f = ln.split("|")
        if len(f) >= 3 and f[0].strip() == kernel:
            try:
                return float(f[-1])
            except ValueError:
                return None
    return None


def copy_calibration(mib=256, reps=5):
    """A plain device-to-device copy of known size (for scaling the
    FETCH_SIZE / WRITE_SIZE counters of rocprofv3 --pmc runs)."""
    import torch
    a = torch.ones(mib * 1024 * 1024 // 4, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(reps):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_read": a.numel() * 4, "copy_bytes_written": b.numel() * 4, "reps": reps}),
          flush=True)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def workload(name):
    if name == "nacl64k":
        s = systems.nacl(ncells=20, kvec=(128, 128, 128), padding=0.5)
        desc = ("molten NaCl 64,000 ions, BHM (Tosi-Fumi) + SPME, rc 12 A, alpha 0.28/A, mesh 128^3, "
                "order 8, padding 0.5 A, NVE dt 2 fs")
    elif name == "nacl1728k":
        s = systems.nacl(ncells=60, kvec=(360, 360, 360), padding=0.5)
        desc = ("molten NaCl 1,728,000 ions, BHM (Tosi-Fumi) + SPME, rc 12 A, alpha 0.28/A, mesh 360^3, "
                "order 8, padding 0.5 A, NVE dt 2 fs")
    elif name == "argon4k":
        s = systems.argon()
        desc = "LJ argon 4,000 atoms, rc 8.5 A, padding 0.5 A, NVE dt 5 fs"
    else:
        raise SystemExit("unknown workload %s" % name)
    return s, desc


def ensemble_cpu_baseline(syss, steps=6):
    """The compiled reference, serial on 1 core, on a sample of the ensemble
    (the smallest and largest argon and NaCl systems): atom-steps/s over the
    sample's steps 2..steps."""
    argon = sorted([s for s in syss if not s.electrostatics], key=lambda s: s.natms)
    nacl = sorted([s for s in syss if s.electrostatics], key=lambda s: s.natms)
    sample = [x for x in (argon[:1] + argon[-1:] + nacl[:1] + nacl[-1:])]
    tot_as, tot_t, names = 0.0, 0.0, []
    for s in sample:
        r = cpu_baseline(s, steps=steps, compare=False)
        if r is None:
            return None
        tot_as += s.natms * (steps - 1)
        tot_t += s.natms * (steps - 1) / r["value"]
        names.append("%s %d" % ("NaCl" if s.electrostatics else "argon", s.natms))
    return dict(value=tot_as / tot_t, unit="atom-steps/s", cores=1, kind="reference",
                sample="DL_POLY 5.1 serial (oracle/_ref/DLPOLY.Z) on 4 of the ensemble's systems (%s), steps 2..%d "
                       "of a %d-step NVE run each (cpu[s] column of OUTPUT)" % (", ".join(names), steps, steps))


def run_ensemble(args):
    """Mixed-precision ensemble (dlpx_md_run_ensemble over all systems) against
    the per-system FP64 loop (dlpx_md_run one system after another)."""
    syss = systems.ensemble_systems()
    natoms = sum(s.natms for s in syss)
    log("ensemble: %d systems, %d atoms (%d argon, %d NaCl)" % (len(syss), natoms,
        sum(1 for s in syss if not s.electrostatics), sum(1 for s in syss if s.electrostatics)))
    # per-system FP64 loop
    t64 = 0.0
    for s in syss:
        ctx = lib.Context(s, precision=lib.FP64)
        ctx.md_load(s.pos, s.vel)
        ctx.md_run(args.warmup, s.timestep, energies=False)
        t0 = time.perf_counter()
        ctx.md_run(args.steps, s.timestep, energies=False)
        t64 += time.perf_counter() - t0
        ctx.close()
    # mixed ensemble
    ctxs = []
    for s in syss:
        ctx = lib.Context(s, precision=lib.MIXED)
        ctx.md_load(s.pos, s.vel)
        ctxs.append(ctx)
    lib.md_run_ensemble(ctxs, args.warmup)
    t0 = time.perf_counter()
    nreb = lib.md_run_ensemble(ctxs, args.steps)
    tmx = time.perf_counter() - t0
    # a second, separately timed pass of eager steps with per-stage HIP events
    # on the library stream; the pair kernel's roofline over every system
    ctxs[0].set_timing(True)
    lib.md_run_ensemble(ctxs, 21)
    ms = ctxs[0].get_timing()
    ctxs[0].set_timing(False)
    nev = max(ms[7], 1.0)
    flops = 0.0
    for c, s in zip(ctxs, syss):
        rc = max(s.cutoff, s.vdw_cutoff) if s.electrostatics else s.vdw_cutoff
        flops += (60.0 if s.electrostatics else 40.0) * c.count_pairs(rc)
    t_pair = ms[0] / nev * 1e-3
    roof = dict(bound="valu", achieved=flops / t_pair / 1e12, peak=VALU_PEAK[lib.MIXED] / 1e12, unit="TFLOP/s")
    roof["frac"] = roof["achieved"] / roof["peak"]
    roof["kernel"] = BATCH_PAIR_KERNEL + " (one launch over all systems)"
    roof["traffic"] = traffic_from_profile("ensemble256", BATCH_PAIR_KERNEL)
    stages = {nm: ms[k] / nev for k, nm in enumerate(["pair", "spread", "fft", "gather", "step"])}
    stages["list"] = ms[5] / nev
    for c in ctxs:
        c.close()
    rate_mx, rate_64 = natoms * args.steps / tmx, natoms * args.steps / t64
    cpu = None if args.no_cpu_baseline else ensemble_cpu_baseline(syss)
    return {
        "metric": METRIC + " (batched ensemble)",
        "value": rate_mx,
        "unit": "atom-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": tmx / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 (f64 accumulation)",
        "data": "synthetic (seeded generator: %d systems, %d atoms)" % (len(syss), natoms),
        "config": {"workload": "ensemble of 256 systems, half LJ argon and half molten NaCl, 1,000-20,000 "
                               "atoms each, NaCl meshes 32^3-64^3, order 8; all steps of all systems",
                   "systems": len(syss), "atoms": natoms},
        "fp64_per_system_loop": {"value": rate_64, "ms_per_step": t64 / args.steps * 1e3},
        "mixed_ensemble_over_fp64_loop": rate_mx / rate_64,
        "list_rebuilds": int(nreb.sum()),
        "stages_ms_per_step": stages,
        "roofline": roof,
        "cpu_baseline": cpu,
    }


def stage_work(s, npairs_rc, npairs_vdw_only, precision):
    """Algorithmic flops and bytes per step of each stage (per-unit figures:
    pair 60 flop per in-cutoff Coulomb+vdW pair, 40 per vdW-only pair, 60 B per
    atom; spread 600 + 2 S^3 flop and 36 B per atom plus one grid write;
    FFT+influence 2 x 2.5 K^3 log2 K^3 + 10 K^3 flop and 4 grid passes;
    gather 8 S^3 flop and 24 B per atom plus one grid read)."""
    n, S = s.natms, s.nsplines
    K3 = float(np.prod(s.kvec)) if s.electrostatics else 0.0
    gb = 8.0 if precision == lib.FP64 else 4.0   # grid element bytes
    out = {"pair": (60.0 * npairs_rc + 40.0 * npairs_vdw_only, 60.0 * n)}
    if s.electrostatics:
        out["spread"] = ((600.0 + 2 * S ** 3) * n, 36.0 * n + gb * K3)
        out["fft"] = (2 * 2.5 * K3 * math.log2(K3) + 10 * K3, 4 * gb * K3)
        out["gather"] = (8.0 * S ** 3 * n, 24.0 * n + gb * K3)
    return out


def run(s, precision, steps, warmup):
    ctx = lib.Context(s, precision=precision)
    ctx.md_load(s.pos, s.vel)
    if warmup:
        ctx.md_run(warmup, s.timestep, energies=False)
    ctx.set_timing(False)
    t0 = time.perf_counter()
    _, nreb = ctx.md_run(steps, s.timestep, energies=False)
    t1 = time.perf_counter()
    # a second, separately timed pass with per-stage HIP events on the library stream
    ctx.set_timing(True)
    _, nreb_t = ctx.md_run(steps, s.timestep, energies=False)
    ms = ctx.get_timing()
    ctx.set_timing(False)
    ctx.close()
    return dict(seconds=t1 - t0, rebuilds=nreb, stage_ms=ms, timed_rebuilds=int(np.sum(nreb_t)))


def nve_drift(s, precision, steps):
    """Total-energy drift of a resident NVE run from the bench configuration:
    least-squares slope of E_tot = engcpe + engsrp + kinetic, relative to
    |E_tot(0)|, per ns; and the largest relative excursion from E_tot(0)."""
    ctx = lib.Context(s, precision=precision)
    ctx.md_load(s.pos, s.vel)
    e, _ = ctx.md_run(steps, s.timestep)
    ctx.close()
    etot = e[:, 0] + e[:, 2] + e[:, 4]
    t_ps = np.arange(1, steps + 1) * s.timestep
    slope = np.polyfit(t_ps, etot, 1)[0]  # energy units per ps
    return dict(steps=steps, ps=steps * s.timestep, drift_rel_per_ns=float(slope * 1000.0 / abs(etot[0])),
                max_rel_excursion=float(np.abs(etot - etot[0]).max() / abs(etot[0])))


def force_error(s, md_steps=50):
    """Mixed-precision forces/energies against the FP64 path, on the
    configuration reached after md_steps FP64 NVE steps from the starting one
    (a lattice start can have forces that vanish by symmetry)."""
    ctx = lib.Context(s, precision=lib.FP64)
    ctx.md_load(s.pos, s.vel)
    ctx.md_run(md_steps, s.timestep, energies=False)
    pos, _, _ = ctx.md_store()
    ctx.close()
    res = {}
    for prec in (lib.FP64, lib.MIXED):
        ctx = lib.Context(s, precision=prec)
        np_list, _ = ctx.link_cell_pairs(pos)
        res[prec] = ctx.two_body_forces(pos)
        if prec == lib.FP64:
            rc = max(s.cutoff, s.vdw_cutoff)
            np_rc = ctx.count_pairs(rc)
        ctx.close()
    f64, fmx = res[lib.FP64]["frc"], res[lib.MIXED]["frc"]
    err = float(np.sqrt(((fmx - f64) ** 2).sum() / (f64 ** 2).sum()))
    e64 = res[lib.FP64]["engcpe"] + res[lib.FP64]["engsrp"]
    emx = res[lib.MIXED]["engcpe"] + res[lib.MIXED]["engsrp"]
    v64 = res[lib.FP64]["vircpe"] + res[lib.FP64]["virsrp"]
    vmx = res[lib.MIXED]["vircpe"] + res[lib.MIXED]["virsrp"]
    return dict(force_rms_rel=err, energy_rel=abs(emx - e64) / abs(e64), virial_rel=abs(vmx - v64) / abs(v64),
                npairs_rc=np_rc, npairs_list=int(np_list))


def vs_reference(s, steps, revcon):
    """Force RMS relative error of the GPU FP64 and mixed paths against the
    reference's REVCON forces (printed to 10 significant figures) after the
    same number of NVE steps from the same start."""
    out = {"steps": steps, "reference_digits": 10}
    for prec, name in ((lib.FP64, "fp64"), (lib.MIXED, "mixed")):
        ctx = lib.Context(s, precision=prec)
        ctx.md_load(s.pos, s.vel)
        ctx.md_run(steps, s.timestep, energies=False)
        _, _, f = ctx.md_store()
        ctx.close()
        ref = revcon["frc"]
        out[name] = float(np.sqrt(((f - ref) ** 2).sum() / (ref ** 2).sum()))
    return out


def output_cpu(out):
    """Cumulative 'cpu [s]' per step from a DL_POLY OUTPUT text."""
    lines = out.splitlines()
    cpu = {}
    for i, ln in enumerate(lines):
        m = re.match(r"^\s+(\d+)\s+-?\d\.\d+E[+-]\d+", ln)
        if m and i + 2 < len(lines):
            try:
                cpu[int(m.group(1))] = float(lines[i + 2].split()[0])
            except (ValueError, IndexError):
                pass
    return cpu


REF_EXE = os.path.join(ROOT, "oracle", "_ref", "DLPOLY.Z")


def cpu_baseline(s, steps=8, compare=True):
    """The compiled reference (oracle/_ref/DLPOLY.Z, serial, 1 core) on the same
    configuration: cumulative 'cpu [s]' of OUTPUT between step 1 and the last
    step (initialisation and the step-0 force call excluded)."""
    if not os.path.exists(REF_EXE):
        return None
    with tempfile.TemporaryDirectory() as d:
        fileio.write_inputs(d, s, steps=steps, print_every=1)
        t0 = time.time()
        r = subprocess.run([REF_EXE], cwd=d, capture_output=True, text=True, timeout=900)
        wall = time.time() - t0
        out = open(os.path.join(d, "OUTPUT")).read() if os.path.exists(os.path.join(d, "OUTPUT")) else ""
        revcon = fileio.read_revcon(os.path.join(d, "REVCON")) if os.path.exists(os.path.join(d, "REVCON")) else None
    if r.returncode != 0:
        log("reference run failed:", r.stdout[-500:])
        return None
    cpu = output_cpu(out)
    if 1 not in cpu or steps not in cpu:
        return None
    dt = cpu[steps] - cpu[1]
    vs_ref = vs_reference(s, steps, revcon) if (compare and revcon is not None) else None
    return dict(value=s.natms * (steps - 1) / dt, unit="atom-steps/s", cores=1, kind="reference",
                forces_vs_reference=vs_ref,
                sample="DL_POLY 5.1 serial (oracle/_ref/DLPOLY.Z), same %d-ion configuration, steps 2..%d of a "
                       "%d-step NVE run (cpu[s] column of OUTPUT, %.1f s wall incl. set-up)"
                       % (s.natms, steps, steps, wall))


def cpu_baseline_box(s, steps=6, nproc=None):
    """The host's throughput bound without an MPI build: nproc concurrent serial
    runs of the compiled reference on the same configuration, one per core
    (the GPU box's CPU share, at most 16), each rated as in cpu_baseline; the
    figure is their summed atom-steps/s."""
    if not os.path.exists(REF_EXE):
        return None
    nproc = nproc or max(1, min(16, os.cpu_count() or 1))
    dirs = [tempfile.mkdtemp() for _ in range(nproc)]
    try:
        for d in dirs:
            fileio.write_inputs(d, s, steps=steps, print_every=1)
        t0 = time.time()
        procs = [subprocess.Popen([REF_EXE], cwd=d, stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL) for d in dirs]
        rcs = [p.wait(timeout=900) for p in procs]
        wall = time.time() - t0
        rates = []
        for d, rc in zip(dirs, rcs):
            f = os.path.join(d, "OUTPUT")
            cpu = output_cpu(open(f).read()) if rc == 0 and os.path.exists(f) else {}
            if 1 not in cpu or steps not in cpu:
                return None
            rates.append(s.natms * (steps - 1) / (cpu[steps] - cpu[1]))
    finally:
        import shutil
        for d in dirs:
            shutil.rmtree(d, ignore_errors=True)
    return dict(value=float(sum(rates)), unit="atom-steps/s", cores=nproc, kind="reference",
                sample="%d concurrent serial DL_POLY 5.1 runs (oracle/_ref/DLPOLY.Z; no MPI library on the box), one per "
                       "host core, each on the same %d-ion configuration, steps 2..%d of a %d-step NVE run (cpu[s] "
                       "column of OUTPUT); per-run rates %.3g-%.3g atom-steps/s, %.1f s wall"
                       % (nproc, s.natms, steps, steps, min(rates), max(rates), wall))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", default="nacl64k")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-procs", type=int, default=0, help="concurrent reference runs of the per-box CPU bound (0: "
                    "one per host core, at most 16)")
    ap.add_argument("--drift-steps", type=int, default=10000,
                    help="NVE steps of the energy-drift runs (0 skips them)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a 256 MiB device copy (counter calibration for rocprofv3 --pmc)")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration()
        return
    if args.gpus != 1:
        print("bench.py measures one GPU: this build runs the force path on a single MI355X "
              "(multi-GPU sharding is a design note only); use --gpus 1", file=sys.stderr)
        sys.exit(2)
    if args.workload == "ensemble256":
        print(json.dumps(run_ensemble(args)), flush=True)
        return
    s, desc = workload(args.workload)
    log("workload:", desc)

    err = force_error(s)
    log("mixed vs fp64 (after 50 FP64 steps):", err)
    runs = {}
    for prec in (lib.FP64, lib.MIXED):
        runs[prec] = run(s, prec, args.steps, args.warmup)
        log("precision", prec, runs[prec])
    rate = {p: s.natms * args.steps / runs[p]["seconds"] for p in runs}

    # roofline per stage (algorithmic work / measured stage time) for the mixed path
    npairs_rc = err["npairs_rc"]
    work = stage_work(s, npairs_rc if s.electrostatics else 0, 0 if s.electrostatics else npairs_rc, lib.MIXED)
    ms = runs[lib.MIXED]["stage_ms"]
    nev = max(ms[7], 1.0)
    names = ["pair", "spread", "fft", "gather"]
    stages = {}
    tmin_sum = tmeas_sum = 0.0
    for k, nm in enumerate(names):
        if nm not in work:
            continue
        t = ms[k] / nev * 1e-3  # s per launch
        flops, byts = work[nm]
        t_flop, t_byte = flops / VALU_PEAK[lib.MIXED], byts / HBM_PEAK
        tmin = max(t_flop, t_byte)
        stages[nm] = dict(ms=t * 1e3, bound="hbm" if t_byte >= t_flop else "valu",
                          achieved_tflops=flops / t / 1e12, achieved_gbs=byts / t / 1e9, tmin_frac=tmin / t)
        tmin_sum += tmin
        tmeas_sum += t
    # the dominant single kernel (the FFT stage is three launches of two kernels)
    dom = max((k for k in stages if k != "fft"), key=lambda k: stages[k]["ms"])
    d = stages[dom]
    flops, byts = work[dom]
    t = d["ms"] * 1e-3
    if d["bound"] == "hbm":
        roof = dict(bound="hbm", achieved=byts / t / 1e9, peak=HBM_PEAK / 1e9, unit="GB/s")
    else:
        roof = dict(bound="valu", achieved=flops / t / 1e12, peak=VALU_PEAK[lib.MIXED] / 1e12, unit="TFLOP/s")
    roof["frac"] = roof["achieved"] / roof["peak"]
    # (without electrostatics the combined-table kernel does not apply)
    roof["kernel"] = KERNELS[dom] if (s.electrostatics or dom != "pair") else PAIR_KERNEL_VDW
    roof["traffic"] = traffic_from_profile(args.workload, roof["kernel"])
    roof["stage_fraction_sum_tmin_over_tmeasured"] = tmin_sum / tmeas_sum if tmeas_sum else None
    if "fft" in stages:
        tr = [traffic_from_profile(args.workload, k) for k, _ in FFT_KERNELS]
        stages["fft"]["traffic"] = (sum(t * m for t, (_, m) in zip(tr, FFT_KERNELS))
                                    if all(t is not None for t in tr) else None)
    # the list build, per rebuild (28 B per atom read + 4 B per stored entry;
    # HBM-bound), outside the per-step sum above
    t_list = ms[5] / max(runs[lib.MIXED]["timed_rebuilds"], 1) * 1e-3
    if runs[lib.MIXED]["timed_rebuilds"] > 0:
        b_list = 28.0 * s.natms + 4.0 * err["npairs_list"]
        stages["list_per_rebuild"] = dict(ms=t_list * 1e3, bound="hbm", achieved_gbs=b_list / t_list / 1e9,
                                          tmin_frac=b_list / HBM_PEAK / t_list, entries=err["npairs_list"])

    drift = None
    if args.drift_steps > 0:
        drift = {"mixed": nve_drift(s, lib.MIXED, args.drift_steps), "fp64": nve_drift(s, lib.FP64, args.drift_steps)}
        log("nve drift:", drift)
    cpu = None
    if not args.no_cpu_baseline:
        # the box's host cores (concurrent serial reference runs), with the
        # single-core run (and its force comparison with the reference) inside
        single = cpu_baseline(s)
        cpu = cpu_baseline_box(s, nproc=args.cpu_procs or None)
        if cpu is None:
            cpu = single
        elif single is not None:
            cpu["single_core"] = single
    out = {
        "metric": METRIC,
        "value": rate[lib.MIXED],
        "unit": "atom-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": runs[lib.MIXED]["seconds"] / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 (f64 accumulation)",
        "data": "synthetic (seeded generator: lattice + jitter + Maxwell-Boltzmann velocities)",
        "config": {"workload": desc, "natms": s.natms, "kvec": list(s.kvec), "nsplines": s.nsplines,
                   "cutoff": s.cutoff, "padding": s.padding, "alpha": s.alpha, "timestep_ps": s.timestep},
        "fp64": {"value": rate[lib.FP64], "ms_per_step": runs[lib.FP64]["seconds"] / args.steps * 1e3,
                 "list_rebuilds": runs[lib.FP64]["rebuilds"]},
        "mixed_over_fp64": rate[lib.MIXED] / rate[lib.FP64],
        "mixed_error_vs_fp64": dict({k: err[k] for k in ("force_rms_rel", "energy_rel", "virial_rel")},
                                    configuration="after 50 FP64 NVE steps from the start"),
        "list_rebuilds": runs[lib.MIXED]["rebuilds"],
        # list build (binning, sort, permute excluded): ms per rebuild, from the timed pass
        "list_build_ms": {("mixed" if p == lib.MIXED else "fp64"): runs[p]["stage_ms"][5] / max(runs[p]["timed_rebuilds"], 1)
                          for p in runs},
        "stages_mixed": stages,
        "stages_fp64_ms": {nm: runs[lib.FP64]["stage_ms"][k] / max(runs[lib.FP64]["stage_ms"][7], 1.0)
                           for k, nm in enumerate(names) if nm in work},
        "nve_drift": drift,
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
