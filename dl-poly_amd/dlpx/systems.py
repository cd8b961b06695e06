"""Synthetic MD systems for the force path: molten NaCl and Lennard-Jones argon.

Everything is in DL_POLY internal units: length in Angstrom, time in ps, mass
in Dalton, energy in 10 J/mol (``constants.F90:100-116``), charge in e.

A ``System`` carries exactly what DL_POLY's CONFIG/FIELD/CONTROL trio would
carry for the two-body + SPME path: cell (row vectors a, b, c as in
``configuration_type%cell``), positions, velocities, per-atom type index,
per-type mass/charge, the vdW pair table (``vdw.F90`` type keys) and the
cutoffs / Ewald / SPME settings.
"""
import math
from dataclasses import dataclass, field
from typing import List, Optional, Tuple

import numpy as np

# constants.F90:100,104 (internal units)
R4PIE0 = 138935.4835
BOLTZ = 8.31451115e-1

# vdw.F90:64-113 potential keys of the forms the library generates, their
# FIELD keywords (ffield.F90 vdw section) and parameter counts
VDW_12_6 = 1
VDW_LENNARD_JONES = 2
VDW_N_M = 3
VDW_BUCKINGHAM = 4
VDW_BORN_HUGGINS_MEYER = 5
VDW_HYDROGEN_BOND = 6
VDW_N_M_SHIFT = 7           # params e0, n, m, r0, r_trunc
VDW_MORSE = 8
VDW_WCA = 9                 # params eps, sig, d, cut = 2^(1/6) sig + d
VDW_DPD = 10                # params a, rc
VDW_AMOEBA = 11             # params eps, sig (14-7)
VDW_LENNARD_JONES_COHESIVE = 12
VDW_MORSE_12 = 13
VDW_RYDBERG = 14
VDW_ZBL = 15                # params Z1, Z2
VDW_ZBL_SWITCH_MORSE = 16   # params Z1, Z2, rm, ic, e0, r0, k
VDW_ZBL_SWITCH_BUCKINGHAM = 17  # params Z1, Z2, rm, ic, A, rho, C
VDW_LJ_MDF = 18             # params eps, sig, ri (damped from ri to the vdW cutoff)
VDW_BUCKINGHAM_MDF = 19     # params A, rho, C, ri
VDW_126_MDF = 20            # params a, b, ri
VDW_LJF = 21                # params ea, sig^2, rc^2 (ljf_params converts eps, sig, rc)
VDW_SANDERSON = 22          # params A, L, d
VDW_NDPD = 23               # params a, b, n, rc
VDW_SW = 24                 # params eps, A, B, sig, p, q, aa
VDW_NAMES = {VDW_12_6: "12-6", VDW_LENNARD_JONES: "lj", VDW_N_M: "nm", VDW_BUCKINGHAM: "buck",
             VDW_BORN_HUGGINS_MEYER: "bhm", VDW_HYDROGEN_BOND: "hbnd", VDW_MORSE: "mors", VDW_WCA: "wca",
             VDW_DPD: "dpd", VDW_LENNARD_JONES_COHESIVE: "ljc", VDW_MORSE_12: "mstw", VDW_RYDBERG: "ryd",
             VDW_SANDERSON: "sand", VDW_NDPD: "ndpd", VDW_N_M_SHIFT: "snm",
             VDW_AMOEBA: "14-7", VDW_ZBL: "zbl", VDW_ZBL_SWITCH_MORSE: "zbls", VDW_ZBL_SWITCH_BUCKINGHAM: "zblb",
             VDW_LJ_MDF: "mlj", VDW_BUCKINGHAM_MDF: "mbuc", VDW_126_MDF: "m126", VDW_LJF: "ljf", VDW_SW: "sw"}
VDW_NPARAMS = {VDW_12_6: 2, VDW_LENNARD_JONES: 2, VDW_N_M: 4, VDW_BUCKINGHAM: 3, VDW_BORN_HUGGINS_MEYER: 5,
               VDW_HYDROGEN_BOND: 2, VDW_MORSE: 3, VDW_WCA: 4, VDW_DPD: 2, VDW_LENNARD_JONES_COHESIVE: 3,
               VDW_MORSE_12: 4, VDW_RYDBERG: 3, VDW_SANDERSON: 3, VDW_NDPD: 4, VDW_N_M_SHIFT: 5, VDW_AMOEBA: 2,
               VDW_ZBL: 2, VDW_ZBL_SWITCH_MORSE: 7, VDW_ZBL_SWITCH_BUCKINGHAM: 7, VDW_LJ_MDF: 3,
               VDW_BUCKINGHAM_MDF: 4, VDW_126_MDF: 3, VDW_LJF: 3, VDW_SW: 7}


def ljf_params(eps, sig, rc):
    """LJF parameters as the FIELD reader converts them (ffield.F90:3818-3829):
    ea = eps * x/4 * (3/(x-1))^3 with x = (rc/sig)^2, then sig^2 and rc^2."""
    x = (rc / sig) ** 2
    a = 0.25 * x * (3.0 / (x - 1.0)) ** 3
    return (eps * a, sig ** 2, rc ** 2)


def ljf_field_words(params):
    """The FIELD words (eps, sig, rc) of converted LJF parameters."""
    ea, sig2, rc2 = params[:3]
    sig, rc = math.sqrt(sig2), math.sqrt(rc2)
    x = (rc / sig) ** 2
    return (ea / (0.25 * x * (3.0 / (x - 1.0)) ** 3), sig, rc)


def wca_params(eps, sig, d=0.0):
    """WCA parameters as the FIELD reader completes them (ffield.F90:3784-3788):
    sig taken positive, d limited to sig/2 from above, cut = 2^(1/6) sig + d."""
    sig = abs(sig)
    if d > sig / 2.0:
        d = sig / 2.0
    return (eps, sig, d, 2.0 ** (1.0 / 6.0) * sig + d)

# Tosi-Fumi NaCl Born-Huggins-Meyer parameters, DL_POLY form
# A*exp(B*(sigma-r)) - C/r^6 - D/r^8, converted from kJ/mol to internal
# units (x100).  Literature values (Tosi & Fumi 1964, b = 0.338e-19 J,
# rho = 0.317 A, Pauling factors 1.25/1.00/0.75).
TOSI_FUMI = {
    (0, 0): (2544.4, 3.15457, 2.340, 10117.0, 4818.0),
    (1, 0): (2035.5, 3.15457, 2.755, 67448.0, 83708.0),
    (1, 1): (1526.6, 3.15457, 3.170, 698570.0, 1403160.0),
}
NA_MASS, CL_MASS, AR_MASS = 22.98977, 35.453, 39.948


@dataclass
class Molecule:
    """One FIELD molecule type: ``nummols`` copies of its sites (0-based atom
    type per site, in CONFIG order) and the bonds and angles that make
    exclusions (0-based site indices).  This path computes no bonded forces:
    the terms only mark which pairs leave the two-body sum, as
    build_excl_intra does (build_excl.F90:32-1195: a bond excludes its two
    atoms, an angle all three pairs)."""
    name: str
    nummols: int
    sites: List[int]
    bonds: List[Tuple[int, int]] = field(default_factory=list)
    angles: List[Tuple[int, int, int]] = field(default_factory=list)


@dataclass
class System:
    name: str
    cell: np.ndarray                 # (9,) a, b, c row vectors
    pos: np.ndarray                  # (N, 3) float64
    vel: np.ndarray                  # (N, 3) float64
    types: np.ndarray                # (N,) int32, 0-based type index
    type_names: List[str]
    masses: np.ndarray               # (ntypes,)
    charges: np.ndarray              # (ntypes,)
    vdw: List[Tuple[int, int, int, Tuple[float, ...]]]  # (ti, tj, key, params)
    cutoff: float = 12.0             # neigh%cutoff (Coulomb / list)
    padding: float = 0.0             # neigh%padding
    vdw_cutoff: float = 12.0         # vdws%cutoff
    electrostatics: bool = True      # SPME on/off
    alpha: float = 0.28              # ewld%alpha
    kvec: Tuple[int, int, int] = (32, 32, 32)
    nsplines: int = 8
    epsilon: float = 1.0             # electro%eps
    timestep: float = 0.002          # ps
    temperature: float = 0.0
    vdw_force_shift: bool = False    # vdws%l_force_shift (CONTROL vdw_force_shift on)
    vdw_direct: bool = False         # vdws%l_direct (CONTROL vdw_method direct)
    meta: dict = field(default_factory=dict)
    molecules: Optional[List[Molecule]] = None  # None: one single-atom molecule type per atom type
    frozen: Optional[np.ndarray] = None  # (N,) bool, config%lfrzn (FIELD frozen flag); None: none

    @property
    def natms(self):
        return int(self.pos.shape[0])

    @property
    def ntypes(self):
        return len(self.type_names)

    @property
    def q(self):
        return self.charges[self.types].astype(np.float64)

    @property
    def mass(self):
        return self.masses[self.types].astype(np.float64)

    @property
    def volume(self):
        c = self.cell.reshape(3, 3)
        return float(abs(np.linalg.det(c)))

    def exclusions(self):
        """Excluded partners of every atom (sorted 0-based indices), as
        neigh%list_excl holds them, or None when nothing is excluded."""
        if not self.molecules:
            return None
        excl = [set() for _ in range(self.natms)]
        base = 0
        for m in self.molecules:
            pairs = [tuple(b) for b in m.bonds]
            for (a, b, c) in m.angles:
                pairs += [(a, b), (a, c), (b, c)]
            for k in range(m.nummols):
                o = base + k * len(m.sites)
                for (a, b) in pairs:
                    excl[o + a].add(o + b)
                    excl[o + b].add(o + a)
            base += m.nummols * len(m.sites)
        if base != self.natms:
            raise ValueError("molecules describe %d atoms, the system has %d" % (base, self.natms))
        if not any(excl):
            return None
        return [np.array(sorted(e), np.int64) for e in excl]


def _streams(seed, n):
    ss = np.random.SeedSequence(seed)
    return [np.random.Generator(np.random.PCG64(s)) for s in ss.spawn(n)]


def _maxwell(rng, mass, temperature):
    """Maxwell-Boltzmann velocities at T with the centre-of-mass momentum removed."""
    sig = np.sqrt(BOLTZ * temperature / mass)[:, None]
    v = rng.standard_normal((mass.shape[0], 3)) * sig
    p = (mass[:, None] * v).sum(0)
    v -= p / mass.sum()
    return v


def _wrap(pos, L):
    """Fold into [-L/2, L/2) per axis (DL_POLY's origin is the cell centre);
    L is a scalar or a per-axis vector."""
    return pos - L * np.floor(pos / L + 0.5)


FCC = np.array([[0.0, 0.0, 0.0], [0.0, 0.5, 0.5], [0.5, 0.0, 0.5], [0.5, 0.5, 0.0]])


def nacl(ncells=20, a0=5.64, expand=1.08, temperature=1200.0, jitter=0.1, seed=20261202,
         kvec=None, nsplines=8, alpha=0.28, cutoff=12.0, padding=0.5, timestep=0.002,
         jitter_on=True, charges=(1.0, -1.0)):
    """Molten NaCl: ncells^3 rocksalt cells (8 ions each), lattice a0*expand
    (the 8% expansion read as linear), positions jittered by U(-jitter, jitter)
    and Maxwell-Boltzmann velocities at ``temperature``.  Na ions are listed
    before Cl ions, matching a FIELD with molecule types Na then Cl."""
    a = a0 * expand
    nc = np.array([ncells] * 3 if np.isscalar(ncells) else list(ncells))
    Lv = nc * a
    r_jit, r_vel = _streams(seed, 2)
    g = np.stack(np.meshgrid(np.arange(nc[0]), np.arange(nc[1]), np.arange(nc[2]),
                             indexing="ij"), -1).reshape(-1, 3).astype(np.float64)
    na = (g[:, None, :] + FCC[None, :, :]).reshape(-1, 3) * a
    cl = na + np.array([0.5 * a, 0.0, 0.0])
    pos = np.concatenate([na, cl]) - 0.5 * Lv
    n = pos.shape[0]
    if jitter_on and jitter > 0:
        pos = pos + r_jit.uniform(-jitter, jitter, size=(n, 3))
    pos = _wrap(pos, Lv)
    types = np.concatenate([np.zeros(n // 2, np.int32), np.ones(n // 2, np.int32)])
    masses = np.array([NA_MASS, CL_MASS])
    vel = _maxwell(r_vel, masses[types], temperature) if temperature > 0 else np.zeros((n, 3))
    if kvec is None:
        kvec = tuple(smooth235_at_least(int(np.ceil(x))) for x in Lv)
    vdw = [(i, j, VDW_BORN_HUGGINS_MEYER, TOSI_FUMI[(i, j)]) for (i, j) in [(0, 0), (1, 0), (1, 1)]]
    return System(name="nacl_%d" % n, cell=np.diag(Lv).reshape(9).astype(np.float64),
                  pos=pos, vel=vel, types=types, type_names=["Na", "Cl"], masses=masses,
                  charges=np.array(charges, dtype=np.float64), vdw=vdw, cutoff=cutoff, padding=padding,
                  vdw_cutoff=cutoff, electrostatics=True, alpha=alpha, kvec=tuple(kvec),
                  nsplines=nsplines, timestep=timestep, temperature=temperature,
                  meta=dict(seed=seed, a=a, ncells=tuple(int(x) for x in nc), jitter=jitter))


def argon(ncells=10, density=1.374, temperature=120.0, seed=20261202, epsilon=99.6,
          sigma=3.405, cutoff=8.5, padding=0.5, timestep=0.005, jitter=0.0):
    """Lennard-Jones argon on an FCC lattice (4 atoms per cell), cubic box at
    mass density ``density`` g/cm^3, Maxwell-Boltzmann velocities at T.
    epsilon = 0.996 kJ/mol = 99.6 internal units.  ``jitter`` > 0 displaces
    every coordinate by U(-jitter, jitter) (the perfect lattice has zero force)."""
    n = 4 * ncells ** 3
    # number density in A^-3: rho / (m * 1.66053906660e-24 g) * 1e-24 cm^3/A^3
    nd = density / (AR_MASS * 1.66053906660)
    L = (n / nd) ** (1.0 / 3.0)
    a = L / ncells
    r_vel, r_jit = _streams(seed, 2)
    g = np.stack(np.meshgrid(np.arange(ncells), np.arange(ncells), np.arange(ncells),
                             indexing="ij"), -1).reshape(-1, 3).astype(np.float64)
    pos = (g[:, None, :] + FCC[None, :, :] + 0.25).reshape(-1, 3) * a - 0.5 * L
    if jitter > 0:
        pos = pos + r_jit.uniform(-jitter, jitter, size=pos.shape)
    pos = _wrap(pos, L)
    types = np.zeros(n, np.int32)
    masses = np.array([AR_MASS])
    vel = _maxwell(r_vel, masses[types], temperature) if temperature > 0 else np.zeros((n, 3))
    return System(name="argon_%d" % n, cell=np.diag([L, L, L]).reshape(9).astype(np.float64),
                  pos=pos, vel=vel, types=types, type_names=["Ar"], masses=masses,
                  charges=np.array([0.0]), vdw=[(0, 0, VDW_LENNARD_JONES, (epsilon, sigma))],
                  cutoff=cutoff, padding=padding, vdw_cutoff=cutoff, electrostatics=False,
                  alpha=0.0, kvec=(0, 0, 0), nsplines=8, timestep=timestep,
                  temperature=temperature, meta=dict(seed=seed, a=a, ncells=ncells))


# three-species mixture: one vdW form per pair key (internal units, A)
MIX_NAMES = ["Xa", "Xb", "Xc"]
MIX_MASSES = (39.948, 20.18, 83.8)
MIX_VDW = [
    (0, 0, VDW_12_6, (4 * 99.6 * 3.405 ** 12, 4 * 99.6 * 3.405 ** 6)),
    (1, 0, VDW_BUCKINGHAM, (3.2e7, 0.30, 4.28e5)),
    (2, 0, VDW_HYDROGEN_BOND, (8.4e9, 6.3e8)),
    (1, 1, VDW_N_M, (60.0, 12.0, 6.0, 3.1)),
    (2, 1, VDW_MORSE, (80.0, 3.9, 1.6)),
    (2, 2, VDW_LENNARD_JONES, (140.0, 3.6)),
]
# AMOEBA 14-7, Sanderson, nDPD (rc 9.5 A, beyond a 9 A vdW cutoff) and the
# shifted n-m form (truncated at 8.5 A) beside LJ and Buckingham pairs
MIX3_VDW = [
    (0, 0, VDW_AMOEBA, (99.6, 3.9)),
    (1, 0, VDW_SANDERSON, (200.0, 4.0, 1.0)),
    (2, 0, VDW_NDPD, (300.0, 2.0, 3.0, 9.5)),
    (1, 1, VDW_N_M_SHIFT, (120.0, 12.0, 6.0, 3.7, 8.5)),
    (2, 1, VDW_BUCKINGHAM, (3.2e7, 0.30, 4.28e5)),
    (2, 2, VDW_LENNARD_JONES, (140.0, 3.6)),
]
# the other analytic forms: LJ cohesive, Morse + c/r^12, Rydberg, WCA (cut
# 4.2 A, inside the nearest-neighbour shell) and DPD with rc 9.5 A, beyond a
# 9 A vdW cutoff, so the force-shift exception for DPD shows
MIX2_VDW = [
    (0, 0, VDW_LENNARD_JONES_COHESIVE, (99.6, 3.405, 0.8)),
    (1, 0, VDW_MORSE_12, (80.0, 3.9, 1.6, 1.0e7)),
    (2, 0, VDW_RYDBERG, (2.0e5, -1.0e3, 0.4)),
    (1, 1, VDW_WCA, wca_params(120.0, 3.3, 0.5)),
    (2, 1, VDW_DPD, (300.0, 9.5)),
    (2, 2, VDW_LENNARD_JONES, (140.0, 3.6)),
]

# the ZBL family (bare, switched to Morse at 4 A and to Buckingham at 3.8 A,
# near the nearest-neighbour distance so both sides of each switch count),
# LJF (rc 7 A) and the Stillinger-Weber two-body part (cut 5.4 A)
MIX4_VDW = [
    (0, 0, VDW_ZBL, (18.0, 18.0)),
    (1, 0, VDW_ZBL_SWITCH_MORSE, (10.0, 18.0, 4.0, 0.3, 80.0, 3.9, 1.6)),
    (2, 0, VDW_ZBL_SWITCH_BUCKINGHAM, (36.0, 18.0, 3.8, 0.35, 3.2e7, 0.30, 4.28e5)),
    (1, 1, VDW_LJF, ljf_params(120.0, 3.25, 7.0)),
    (2, 1, VDW_SW, (200.0, 7.049556277, 0.6022245584, 3.0, 4.0, 0.0, 1.8)),
    (2, 2, VDW_LENNARD_JONES, (140.0, 3.6)),
]
# the damped (MDF) forms from ri = 7, 6.5 and 7.5 A to the 9 A cutoff, LJF
# with rc 9.5 A (beyond the cutoff, so the direct force-shift exception for
# LJF shows) and a bare ZBL pair
MIX5_VDW = [
    (0, 0, VDW_LJ_MDF, (99.6, 3.405, 7.0)),
    (1, 0, VDW_BUCKINGHAM_MDF, (3.2e7, 0.30, 4.28e5, 6.5)),
    (2, 0, VDW_126_MDF, (4 * 99.6 * 3.405 ** 12, 4 * 99.6 * 3.405 ** 6, 7.5)),
    (1, 1, VDW_LJF, ljf_params(120.0, 3.3, 9.5)),
    (2, 1, VDW_ZBL, (10.0, 36.0)),
    (2, 2, VDW_LENNARD_JONES, (140.0, 3.6)),
]


def mixture(ncells=10, a=5.78015, temperature=100.0, seed=20261202, charges=(0.3, -0.3, 0.0), kvec=(36, 36, 36),
            alpha=0.35, cutoff=9.0, padding=0.5, timestep=0.004, jitter=0.1, vdw=None):
    """Three species on an FCC lattice (4 sites per cell, edge ``a``), the
    species drawn by a seeded permutation of the sites, with a different vdW
    form for each pair (by default 12-6, Buckingham, 12-10, n-m, Morse,
    Lennard-Jones; ``vdw=MIX2_VDW``, ``MIX3_VDW``, ``MIX4_VDW`` and
    ``MIX5_VDW`` give the other forms) and SPME
    Coulomb.  Atoms are listed species by species (the FIELD's molecule
    order)."""
    n = 4 * ncells ** 3
    L = a * ncells
    r_perm, r_jit, r_vel = _streams(seed, 3)
    g = np.stack(np.meshgrid(np.arange(ncells), np.arange(ncells), np.arange(ncells),
                             indexing="ij"), -1).reshape(-1, 3).astype(np.float64)
    sites = (g[:, None, :] + FCC[None, :, :] + 0.25).reshape(-1, 3) * a - 0.5 * L
    sites = sites[r_perm.permutation(n)]
    counts = [n // 3 + (1 if k < n % 3 else 0) for k in range(3)]
    types = np.repeat(np.arange(3, dtype=np.int32), counts)
    pos = _wrap(sites + r_jit.uniform(-jitter, jitter, size=sites.shape), L)
    masses = np.array(MIX_MASSES)
    vel = _maxwell(r_vel, masses[types], temperature) if temperature > 0 else np.zeros((n, 3))
    return System(name="mix_%d" % n, cell=np.diag([L, L, L]).reshape(9).astype(np.float64), pos=pos, vel=vel,
                  types=types, type_names=list(MIX_NAMES), masses=masses, charges=np.array(charges, np.float64),
                  vdw=list(MIX_VDW if vdw is None else vdw), cutoff=cutoff, padding=padding, vdw_cutoff=cutoff,
                  electrostatics=True,
                  alpha=alpha, kvec=tuple(kvec), nsplines=8, timestep=timestep, temperature=temperature,
                  meta=dict(seed=seed, a=a, ncells=ncells, jitter=jitter))


# three-site polar molecule (SPC-like charges and O-O Lennard-Jones, with
# small H repulsions), internal units; the four-site variant adds a site "Dw"
# 0.02 A from Ow that carries most of the oxygen charge (a Drude-like pair,
# closer than 0.01/alpha, so its exclusion correction takes the series branch)
MOL_NAMES = ["Ow", "Hw", "Dw"]
MOL_MASSES = (15.9994, 1.008, 0.4)
MOL_VDW = [
    (0, 0, VDW_LENNARD_JONES, (65.02, 3.166)),
    (1, 0, VDW_LENNARD_JONES, (15.0, 2.0)),
    (1, 1, VDW_LENNARD_JONES, (5.0, 1.2)),
]


def _rotations(rng, n):
    """n uniformly random rotation matrices (from unit quaternions)."""
    q = rng.standard_normal((n, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    w, x, y, z = q.T
    return np.stack([np.stack([1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)], -1),
                     np.stack([2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)], -1),
                     np.stack([2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)], -1)], 1)


def molecular(ncells=6, a=4.93, temperature=300.0, seed=20261202, drude=False, kvec=(32, 32, 32), alpha=0.35,
              cutoff=9.0, padding=0.5, timestep=0.001, bond=1.0, angle=109.47, drude_sep=0.02):
    """Polar triatomic molecules (Ow, Hw, Hw) centred on an FCC lattice (4 per
    cell, edge ``a``) with random orientations, bond length ``bond`` and angle
    ``angle``; with ``drude`` a fourth site Dw sits ``drude_sep`` from Ow.
    Two bonds (and the Dw bond) and the H-O-H angle (and the H-O-Dw angles)
    exclude every intramolecular pair, so the molecules carry no intramolecular
    two-body terms and the reciprocal sum's share of them is removed by the
    Ewald exclusion correction."""
    nmol = 4 * ncells ** 3
    L = a * ncells
    r_rot, r_vel, r_dir = _streams(seed, 3)
    g = np.stack(np.meshgrid(np.arange(ncells), np.arange(ncells), np.arange(ncells),
                             indexing="ij"), -1).reshape(-1, 3).astype(np.float64)
    centres = (g[:, None, :] + FCC[None, :, :] + 0.25).reshape(-1, 3) * a - 0.5 * L
    th = np.radians(angle) / 2
    local = [np.zeros(3), bond * np.array([np.sin(th), 0.0, np.cos(th)]),
             bond * np.array([-np.sin(th), 0.0, np.cos(th)])]
    sites, charges = [0, 1, 1], (-0.82, 0.41, 0.0)
    bonds, angles = [(0, 1), (0, 2)], [(1, 0, 2)]
    if drude:
        local.append(drude_sep * np.array([0.0, 0.0, -1.0]))
        sites, charges = [0, 1, 1, 2], (1.0, 0.41, -1.82)
        bonds.append((0, 3))
        angles += [(1, 0, 3), (2, 0, 3)]
    rot = _rotations(r_rot, nmol)
    pos = centres[:, None, :] + np.einsum("mab,sb->msa", rot, np.array(local))
    pos = _wrap(pos.reshape(-1, 3), L)
    types = np.tile(np.array(sites, np.int32), nmol)
    ntypes = 3 if drude else 2
    masses = np.array(MOL_MASSES[:ntypes])
    if drude:
        masses[0] -= MOL_MASSES[2]
    vel = _maxwell(r_vel, masses[types], temperature) if temperature > 0 else np.zeros_like(pos)
    mol = Molecule(name="polar4" if drude else "polar3", nummols=nmol, sites=sites, bonds=bonds, angles=angles)
    return System(name="mol%d_%d" % (len(sites), pos.shape[0]), cell=np.diag([L, L, L]).reshape(9).astype(np.float64),
                  pos=pos, vel=vel, types=types, type_names=MOL_NAMES[:ntypes], masses=masses,
                  charges=np.array(charges[:ntypes], np.float64), vdw=list(MOL_VDW), cutoff=cutoff,
                  padding=padding, vdw_cutoff=cutoff, electrostatics=True, alpha=alpha, kvec=tuple(kvec),
                  nsplines=8, timestep=timestep, temperature=temperature,
                  meta=dict(seed=seed, a=a, ncells=ncells, drude=drude), molecules=[mol])


def frozen_slab(s, zmax):
    """The same configuration with the atoms below z = zmax frozen
    (config%lfrzn): each frozen species becomes its own atom type (name +
    "f", same mass, charge and vdW parameters), since a FIELD flags frozen
    sites per molecule type, and the atoms are regrouped by type."""
    if s.molecules:
        raise ValueError("frozen_slab works on single-atom molecule layouts")
    nt = s.ntypes
    frz = s.pos[:, 2] < zmax
    types = np.where(frz, s.types + nt, s.types).astype(np.int32)
    order = np.argsort(types, kind="stable")
    species = list(range(nt)) * 2                      # type -> original species
    par = {(max(a, b), min(a, b)): (k, p) for (a, b, k, p) in s.vdw}
    vdw = []
    for i in range(2 * nt):
        for j in range(i + 1):
            key = (max(species[i], species[j]), min(species[i], species[j]))
            if key in par:
                vdw.append((i, j, par[key][0], par[key][1]))
    return System(name=s.name + "_frz", cell=s.cell.copy(), pos=s.pos[order], vel=s.vel[order], types=types[order],
                  type_names=list(s.type_names) + [n + "f" for n in s.type_names],
                  masses=np.concatenate([s.masses, s.masses]), charges=np.concatenate([s.charges, s.charges]),
                  vdw=vdw, cutoff=s.cutoff, padding=s.padding, vdw_cutoff=s.vdw_cutoff,
                  electrostatics=s.electrostatics, alpha=s.alpha, kvec=s.kvec, nsplines=s.nsplines,
                  epsilon=s.epsilon, timestep=s.timestep, temperature=s.temperature,
                  vdw_force_shift=s.vdw_force_shift, vdw_direct=s.vdw_direct, meta=dict(s.meta, frozen_below=zmax),
                  frozen=frz[order])


def molecules_to_json(mols):
    import json
    if not mols:
        return ""
    return json.dumps([dict(name=m.name, nummols=m.nummols, sites=list(m.sites), bonds=[list(b) for b in m.bonds],
                            angles=[list(t) for t in m.angles]) for m in mols])


def molecules_from_json(text):
    import json
    if not text:
        return None
    return [Molecule(name=d["name"], nummols=int(d["nummols"]), sites=[int(x) for x in d["sites"]],
                     bonds=[tuple(b) for b in d["bonds"]], angles=[tuple(t) for t in d["angles"]])
            for d in json.loads(text)]


def smooth235_at_least(n):
    """Smallest 2^a 3^b 5^c >= n (the FFT sizes parallel_fft.F90:2031-2054 accepts)."""
    m = max(1, n)
    while True:
        k = m
        for p in (2, 3, 5):
            while k % p == 0:
                k //= p
        if k == 1:
            return m
        m += 1


def pair_key(ti, tj):
    """vdw.F90:1875-1881 key for 1-based types: max*(max-1)/2 + min (returned 0-based)."""
    a, b = ti + 1, tj + 1
    hi, lo = max(a, b), min(a, b)
    return hi * (hi - 1) // 2 + lo - 1


def ensemble_systems(nsys=256, seed=20261202):
    """The batched-ensemble workload: nsys systems drawn from a seed, half LJ
    argon and half molten NaCl (shuffled), sizes uniform in 1,000-20,000 atoms
    rounded to whole lattices (4 m^3 argon, 8 m^3 NaCl), cubic boxes at the
    densities of the single-system workloads, NaCl meshes the smallest
    2^a 3^b 5^c size of at least one point per A, clamped to 32..64."""
    rng = np.random.default_rng(seed)
    kinds = ["argon"] * (nsys // 2) + ["nacl"] * (nsys - nsys // 2)
    rng.shuffle(kinds)
    out = []
    for i, kind in enumerate(kinds):
        n = int(rng.integers(1000, 20001))
        if kind == "argon":
            m = max(4, int(round((n / 4.0) ** (1.0 / 3.0))))
            out.append(argon(ncells=m, seed=seed + 1 + i))
        else:
            m = max(5, int(round((n / 8.0) ** (1.0 / 3.0))))
            L = m * 5.64 * 1.08
            K = min(64, max(32, smooth235_at_least(int(math.ceil(L)))))
            out.append(nacl(ncells=m, kvec=(K, K, K), seed=seed + 1 + i))
    return out
