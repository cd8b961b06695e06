"""Load the golden fixtures written by tests/golden/make_fixtures.py."""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
for p in (os.path.join(ROOT, "dl-poly_amd"), os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)

from dlpx import systems  # noqa: E402
from dlpx.fileio import STP_ENGCPE, STP_ENGSRP, STP_VIRCPE, STP_VIRSRP  # noqa: E402

NAMES = sorted(f[:-4] for f in os.listdir(GOLDEN) if f.endswith(".npz"))


def load(name):
    z = np.load(os.path.join(GOLDEN, name + ".npz"))
    st = z["settings"]
    vdw = [(int(t[0]), int(t[1]), int(k), tuple(p[:systems.VDW_NPARAMS[int(k)]]))
           for t, k, p in zip(z["vdw_types"], z["vdw_kind"], z["vdw_params"])]
    s = systems.System(name=name, cell=z["cell"], pos=z["pos"], vel=z["vel"], types=z["types"],
                       type_names=[str(x) for x in z["type_names"]], masses=z["masses"],
                       charges=z["charges"], vdw=vdw, cutoff=float(st[0]), padding=float(st[1]),
                       vdw_cutoff=float(st[2]), electrostatics=bool(z["electrostatics"]),
                       alpha=float(st[3]), kvec=tuple(int(k) for k in z["kvec"]),
                       nsplines=int(z["nsplines"]), epsilon=float(st[4]), timestep=float(st[5]),
                       temperature=float(st[6]),
                       vdw_force_shift=bool(z["vdw_force_shift"]) if "vdw_force_shift" in z.files else False,
                       vdw_direct=bool(z["vdw_direct"]) if "vdw_direct" in z.files else False,
                       molecules=systems.molecules_from_json(str(z["molecules"])) if "molecules" in z.files else None,
                       frozen=(z["frozen"].astype(bool) if "frozen" in z.files and z["frozen"].size else None))
    ref = dict(frc=z["ref_frc"], pos=z["ref_pos"], vel=z["ref_vel"], stress=z["ref_stress"],
               stpval=z["ref_stpval"], steps=int(z["steps"]),
               engsrp=z["ref_stpval"][STP_ENGSRP], engcpe=z["ref_stpval"][STP_ENGCPE],
               virsrp=z["ref_stpval"][STP_VIRSRP], vircpe=z["ref_stpval"][STP_VIRCPE])
    # the reference's step-0 state at full precision (static cases; see
    # tests/golden/ref_dump.F90): two_body_forces' own forces, its energies,
    # stress and vdW long-range corrections, and its vdW tables
    if "ref_frc64" in z.files:
        d = z["ref_dump_stats"]
        ref.update(frc64=z["ref_frc64"], dump_engcpe=d[0], dump_vircpe=d[1], dump_engsrp=d[2],
                   dump_virsrp=d[3], dump_stress=d[4:13], elrc=d[13], vlrc=d[14])
    if "ref_tab_pot" in z.files:
        ref.update(tab_pot=z["ref_tab_pot"], tab_frc=z["ref_tab_frc"], tab_list=z["ref_tab_list"])
    return s, ref


def reference_tables(s, ref):
    """The reference's own vdW tables laid out per type-pair key, as
    dlpx_set_vdw_tables takes them: (max_grid, pot, frc, ltp), with ltp the
    vdW form of each key (-1 for a pair without a potential)."""
    nkey = s.ntypes * (s.ntypes + 1) // 2
    m = ref["tab_pot"].shape[1]
    pot, frc, ltp = np.zeros((nkey, m)), np.zeros((nkey, m)), np.full(nkey, -1, np.int32)
    for (ti, tj, kind, _) in s.vdw:
        a, b = ti + 1, tj + 1
        hi, lo = max(a, b), min(a, b)
        key = hi * (hi - 1) // 2 + lo - 1
        k = int(ref["tab_list"][key]) - 1
        pot[key], frc[key], ltp[key] = ref["tab_pot"][k], ref["tab_frc"][k], kind
    return m - 1, pot, frc, ltp


def printed_digits_bound(ref, sig=10):
    """Half a unit in the last of `sig` printed significant figures (g20.10)."""
    a = np.abs(ref)
    e = np.floor(np.log10(np.where(a > 0, a, 1.0)))
    return 0.5 * 10.0 ** (e - (sig - 1))


def heat_flux(s, pp_energy, pp_stress):
    """statistics.F90:2643-2665 (internal energy units): (sum (e_i + m v^2/2) v
    + sum S_i v) / V, S_i the 3x3 of pp_stress(:, i) (column-major)."""
    from dlpx import runner
    v, m = runner.start_velocities(s), s.mass  # as the reference starts the run (frozen at rest, no COM drift)
    ke = 0.5 * m * (v * v).sum(1)
    ev = ((pp_energy + ke)[:, None] * v).sum(0)
    S = pp_stress.reshape(-1, 3, 3).transpose(0, 2, 1)  # S[i, a, b] = pp_stress[i, a + 3b]
    return (ev + np.einsum("iab,ib->a", S, v)) / s.volume


def frozen_zeroed(s, frc):
    """Forces as the reference reports them: freeze_atoms zeroes those of
    frozen atoms after the force evaluation (configuration.F90:3240-3258)."""
    if s.frozen is None:
        return frc
    f = np.array(frc, copy=True)
    f[s.frozen] = 0.0
    return f
