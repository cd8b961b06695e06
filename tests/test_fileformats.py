"""DL_POLY file formats on the host side (CPU): the CONFIG/FIELD/CONTROL
readers against the writers and hand-written inputs, the REVCON and STATIS
writers against their readers, the reference harness's comparison rule, and
the STATIS quantities of one step against the reference's own STATIS file
(tests/golden/statis/, written by DLPOLY.Z for the same inputs)."""
import os
import tempfile

import numpy as np
import pytest

import fixtures
import oracle
from dlpx import fileio, runner, systems

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "statis")


@pytest.mark.parametrize("make", [lambda: systems.nacl(ncells=3, kvec=(20, 24, 30), cutoff=8.5, padding=0.3),
                                  lambda: systems.argon(ncells=4)])
def test_inputs_round_trip_through_readers(make):
    s = make()
    with tempfile.TemporaryDirectory() as d:
        fileio.write_inputs(d, s, steps=7, print_every=3)
        t = fileio.system_from_files(d)
    assert np.array_equal(t.pos, s.pos) and np.array_equal(t.vel, s.vel)
    assert np.array_equal(t.cell, s.cell) and np.array_equal(t.types, s.types)
    assert t.type_names == s.type_names and t.vdw == s.vdw
    assert np.array_equal(t.masses, s.masses) and np.array_equal(t.charges, s.charges)
    assert (t.cutoff, t.padding, t.vdw_cutoff, t.timestep) == (s.cutoff, s.padding, s.vdw_cutoff, s.timestep)
    assert t.electrostatics == s.electrostatics
    if s.electrostatics:
        assert (t.alpha, tuple(t.kvec), t.nsplines, t.epsilon) == (s.alpha, tuple(s.kvec), s.nsplines, s.epsilon)
    assert t.meta["steps"] == 7 and t.meta["stats_frequency"] == 3


FIELD_KJ = """two-site fluid
units kj
molecules 2
pairs
nummols 3
atoms 2
A  12.0  0.5
B  16.0 -0.5
finish
ion
nummols 2
atoms 1
C  40.0  0.0  2
finish
vdw 3
A  A  lj   0.25  3.0
A  B  bhm  10.0  3.0  2.5  50.0  60.0
C  C  lj   1.0   3.4
close
"""


def test_field_reader_units_and_repeats():
    with tempfile.TemporaryDirectory() as d:
        p = os.path.join(d, "FIELD")
        open(p, "w").write(FIELD_KJ)
        f = fileio.read_field(p)
    assert f["units"] == "kj"
    (m1, n1, a1), (m2, n2, a2) = f["molecules"]
    assert (m1, n1, [a[0] for a in a1]) == ("pairs", 3, ["A", "B"])
    assert (m2, n2, a2) == ("ion", 2, [("C", 40.0, 0.0), ("C", 40.0, 0.0)])  # repeat count 2
    lj, bhm, lj2 = f["vdw"]
    # energies scaled to internal units (10 J/mol): kJ/mol x 100; lengths kept
    assert lj == ("A", "A", 2, (25.0, 3.0))
    assert bhm == ("A", "B", 5, (1000.0, 3.0, 2.5, 5000.0, 6000.0))
    assert lj2 == ("C", "C", 2, (100.0, 3.4))


def test_field_reader_converts_the_zbl_mdf_ljf_and_sw_forms():
    """ffield.F90:3760-3839 in kJ/mol: parpot(1) scaled except for the ZBL
    forms (a charge); zbls parpot(5), zblb parpot(5) and (7), mbuc (1) and
    (3), m126 (1) and (2), sw (1); LJF's eps times x/4 (3/(x-1))^3 with
    x = (rc/sig)^2, then sig^2 and rc^2."""
    recs = ["A  C  zbl  14.0 8.0", "B  C  zbls 14.0 8.0 1.5 0.2 2.0 3.0 1.5",
            "A  B  zblb 14.0 8.0 1.5 0.2 3.0 0.3 4.0", "A  A  mlj  0.5 3.0 6.0", "B  A  mbuc 3.0 0.3 4.0 6.5",
            "C  B  m126 5.0 6.0 7.0", "C  A  ljf  2.0 2.0 5.0", "B  B  sw   2.0 7.0 0.6 2.1 4.0 0.0 1.8"]
    head = FIELD_KJ[:FIELD_KJ.index("vdw 3")]
    text = head + "vdw %d\n" % len(recs) + "\n".join(recs) + "\nclose\n"
    with tempfile.TemporaryDirectory() as d:
        p = os.path.join(d, "FIELD")
        open(p, "w").write(text)
        f = fileio.read_field(p)
    got = {(a, b): (k, prm) for (a, b, k, prm) in f["vdw"]}
    assert got[("A", "C")] == (15, (14.0, 8.0))
    assert got[("B", "C")] == (16, (14.0, 8.0, 1.5, 0.2, 200.0, 3.0, 1.5))
    assert got[("A", "B")] == (17, (14.0, 8.0, 1.5, 0.2, 300.0, 0.3, 400.0))
    assert got[("A", "A")] == (18, (50.0, 3.0, 6.0))
    assert got[("B", "A")] == (19, (300.0, 0.3, 400.0, 6.5))
    assert got[("C", "B")] == (20, (500.0, 600.0, 7.0))
    x = (5.0 / 2.0) ** 2
    k, (ea, s2, rc2) = got[("C", "A")]
    assert k == 21 and (s2, rc2) == (4.0, 25.0) and abs(ea - 200.0 * 0.25 * x * (3.0 / (x - 1.0)) ** 3) <= 1e-13 * ea
    assert got[("B", "B")] == (24, (200.0, 7.0, 0.6, 2.1, 4.0, 0.0, 1.8))


def test_field_reader_rejects_buckingham_with_zero_rho():
    """ffield.F90:3745-3751: rho = 0 with A nonzero is DL_POLY's error 467."""
    text = FIELD_KJ.replace("C  C  lj   1.0   3.4", "C  C  buck 2.0 0.0 5.0")
    with tempfile.TemporaryDirectory() as d:
        p = os.path.join(d, "FIELD")
        open(p, "w").write(text)
        with pytest.raises(ValueError, match="467"):
            fileio.read_field(p)


@pytest.mark.parametrize("bad", ["bonds 1\nharm 1 2 100.0 1.0\n", "bonds 1\nmors 1 2 0.0 1.0 2.0\n",
                                 "constraints 1\n1 2 1.0\n", None])
def test_field_reader_rejects_what_the_path_lacks(bad):
    text = FIELD_KJ
    if bad:
        text = text.replace("finish\nion", bad + "finish\nion", 1)
    else:
        text = text.replace("C  C  lj   1.0   3.4", "C  C  tab")  # TABLE-file potentials
    with tempfile.TemporaryDirectory() as d:
        p = os.path.join(d, "FIELD")
        open(p, "w").write(text)
        with pytest.raises(ValueError):
            fileio.read_field(p)


@pytest.mark.parametrize("rec,want", [
    # ffield.F90:3784-3788: sig made positive, d limited to sig/2 from above,
    # cut = 2^(1/6) sig + d; energies in kJ/mol x 100
    ("wca  1.0 -3.0 2.0", (100.0, 3.0, 1.5, 2 ** (1 / 6) * 3.0 + 1.5)),
    ("wca  1.0 3.0 -2.0", (100.0, 3.0, -2.0, 2 ** (1 / 6) * 3.0 - 2.0)),
    ("wca  1.0 3.0", (100.0, 3.0, 0.0, 2 ** (1 / 6) * 3.0)),
    ("dpd  2.0 4.0", (200.0, 4.0)),
    ("ljc  1.0 3.4 0.5", (100.0, 3.4, 0.5)),
    ("mstw 1.0 3.4 1.5 2.0", (100.0, 3.4, 1.5, 200.0)),
    ("ryd  1.0 2.0 0.3", (100.0, 200.0, 0.3)),
    ("sand 2.0 4.0 1.0", (200.0, 4.0, 1.0)),
    ("ndpd 3.0 2.0 3.0 9.5", (300.0, 2.0, 3.0, 9.5)),
    ("snm  1.0 12.0 6.0 3.7 8.5", (100.0, 12.0, 6.0, 3.7, 8.5)),
    ("14-7 1.0 3.9", (100.0, 3.9)),
    # ffield.F90:3745-3751: rho = 0 with A = 0 (pure dispersion) reads as rho = 1
    ("buck 0.0 0.0 5.0", (0.0, 1.0, 500.0)),
    # words past the form's own parameters are read and ignored
    ("bhm  1.0 3.0 2.5 5.0 6.0 7.0", (100.0, 3.0, 2.5, 500.0, 600.0)),
    ("snm  1.0 12.0 6.0 3.7 8.5 9.9", (100.0, 12.0, 6.0, 3.7, 8.5)),
])
def test_field_reader_completes_other_forms(rec, want):
    text = FIELD_KJ.replace("C  C  lj   1.0   3.4", "C  C  " + rec)
    with tempfile.TemporaryDirectory() as d:
        p = os.path.join(d, "FIELD")
        open(p, "w").write(text)
        got = [v for v in fileio.read_field(p)["vdw"] if v[0] == "C" and v[1] == "C"][0]
    assert got[3] == pytest.approx(want, rel=1e-15, abs=0)


FIELD_MOL = """three-site molecules
units internal
molecules 1
polar
nummols 2
atoms 3
Ow 15.9994 -0.82
Hw 1.008 0.41 2
bonds 2
harm 1 2 0.0 1.0
-hrm 1 3 0.0 1.0
angles 1
harm 2 1 3 0.0 109.47
finish
vdw 1
Ow Ow lj 65.02 3.166
close
"""


def test_field_zero_strength_bonds_give_exclusions():
    """bonds and angles mark excluded pairs (build_excl.F90:125-200): a bond
    its two atoms, an angle all three pairs; a restrained term (-hrm)
    excludes nothing."""
    with tempfile.TemporaryDirectory() as d:
        p = os.path.join(d, "FIELD")
        open(p, "w").write(FIELD_MOL)
        f = fileio.read_field(p)
    assert f["topology"] == [([(0, 1)], [(1, 0, 2)])]
    s = systems.molecular(ncells=1)
    s.molecules = [systems.Molecule(name="polar", nummols=2, sites=[0, 1, 1], bonds=[(0, 1)], angles=[(1, 0, 2)])]
    s.pos, s.vel, s.types = s.pos[:6], s.vel[:6], s.types[:6]
    ex = s.exclusions()
    assert [list(e) for e in ex] == [[1, 2], [0, 2], [0, 1], [4, 5], [3, 5], [3, 4]]


def test_frozen_flags_round_trip_and_start_velocities():
    """FIELD frozen flags (``name mass charge repeat frozen``) per atom type,
    and the run's starting velocities: frozen atoms at rest and no
    centre-of-mass drift among the others (temperature.F90:579-620)."""
    s = systems.frozen_slab(systems.nacl(ncells=3, kvec=(24, 24, 24), cutoff=8.0, padding=0.3), -2.0)
    with tempfile.TemporaryDirectory() as d:
        fileio.write_inputs(d, s, steps=2)
        t = fileio.system_from_files(d)
    assert np.array_equal(t.frozen, s.frozen) and t.type_names == ["Na", "Cl", "Naf", "Clf"]
    v = runner.start_velocities(s)
    assert np.all(v[s.frozen] == 0.0)
    p = (s.mass[:, None] * v).sum(0)
    assert np.abs(p).max() <= 1e-12 * (s.mass[:, None] * np.abs(v)).sum()


def test_molecular_inputs_round_trip():
    s = systems.molecular(ncells=2, drude=True)
    with tempfile.TemporaryDirectory() as d:
        fileio.write_inputs(d, s, steps=3)
        t = fileio.system_from_files(d)
    assert np.array_equal(t.types, s.types) and t.type_names == s.type_names
    assert [(m.nummols, m.sites, m.bonds, m.angles) for m in t.molecules] == \
        [(m.nummols, m.sites, m.bonds, m.angles) for m in s.molecules]
    assert all(np.array_equal(a, b) for a, b in zip(t.exclusions(), s.exclusions()))


def test_control_reader_units():
    with tempfile.TemporaryDirectory() as d:
        p = os.path.join(d, "CONTROL")
        open(p, "w").write("title t\nensemble nve\ntimestep 2.0 fs\ntime_run 50 steps\ncutoff 1.2 nm\n"
                           "padding 0.5 ang\ncoul_method spme\nspme_alpha 0.3 ang^-1\nspme_kvec [36 40 45]\n"
                           "stats_frequency 5 steps\nio_file_output OUTPUT.x\n")
        c = fileio.read_control(p)
    assert c["timestep"] == pytest.approx(0.002) and c["steps"] == 50 and c["stats_frequency"] == 5
    assert c["cutoff"] == pytest.approx(12.0) and c["vdw_cutoff"] == pytest.approx(12.0)
    assert c["kvec"] == (36, 40, 45) and c["coul_method"] == "spme"
    assert c["other"]["io_file_output"] == "OUTPUT.x"


def test_revcon_writer_round_trip():
    s = systems.nacl(ncells=2, kvec=(16, 16, 16))
    rng = np.random.default_rng(3)
    frc = rng.normal(scale=1e4, size=(s.natms, 3))
    with tempfile.TemporaryDirectory() as d:
        p = os.path.join(d, "REVCON")
        fileio.write_revcon(p, s.name, s.cell, [s.type_names[t] for t in s.types], s.pos, s.vel, frc,
                            nstep=10, tstep=0.002, time=0.02)
        lines = open(p).read().splitlines()
        assert lines[1].split()[:4] == ["2", "1", str(s.natms), "10"]
        r = fileio.read_config(p)
        rv = fileio.read_revcon(p)
    # 3g20.10 records: 10 significant digits
    for got, want in ((r["pos"], s.pos), (r["vel"], s.vel), (r["frc"], frc), (rv["frc"], frc)):
        assert np.allclose(got, want, rtol=1e-9, atol=0)
    assert r["names"] == [s.type_names[t] for t in s.types]


def test_statis_writer_reader_round_trip():
    vals = np.array([-3.6e7, 1218.06, 0.0, 1.5e-3, -2.0, 90.0, 3.14159e10])
    with tempfile.TemporaryDirectory() as d:
        p = os.path.join(d, "STATIS")
        with open(p, "w") as f:
            f.write(fileio.statis_header("case"))
            f.write(fileio.statis_record(0, 0.0, vals))
            f.write(fileio.statis_record(5, 0.01, vals * 2))
        recs = fileio.read_statis(p)
    assert sorted(recs) == [0, 5] and recs[5][0] == pytest.approx(0.01)
    assert np.allclose(recs[0][1], vals, rtol=1e-6) and np.allclose(recs[5][1], 2 * vals, rtol=1e-6)


def test_statis_equal_is_the_harness_rule():
    # utils/beetest:39-48: both zero pass; zero expected -> absolute; else relative
    assert fileio.statis_equal(0.0, 0.0, 1e-5)
    assert fileio.statis_equal(5e-6, 0.0, 1e-5) and not fileio.statis_equal(2e-5, 0.0, 1e-5)
    assert fileio.statis_equal(1.000009, 1.0, 1e-5) and not fileio.statis_equal(1.00002, 1.0, 1e-5)
    assert not fileio.statis_equal(0.0, 1e-9, 1e-5)


@pytest.mark.parametrize("name", ["nacl_1000_k32", "argon_4000_nve20", "nacl_1000_nve10", "mix_4000_nve10",
                                  "mol3_2592_nve10", "nacl_1000_frz_nve10"])
def test_statis_step0_quantities_match_reference_file(name):
    """The STATIS assembly of one step (temperature, enthalpy, pressure,
    pressure tensor with the kinetic stress, cell angles, energies and
    virials), fed the oracle's step-0 two-body output, reproduces the
    reference's step-0 record in every column at the harness's 1e-5."""
    s, _ = fixtures.load(name)
    ref = fileio.read_statis(os.path.join(GOLDEN, name + ".STATIS"))
    o = oracle.two_body(s)
    vals = runner.statis_values(s, o, runner.start_velocities(s), np.zeros(s.ntypes))
    t0, want = ref[0]
    assert t0 == 0.0 and len(vals) == len(want)
    bad = [(k + 1, vals[k], want[k]) for k in range(len(want)) if not fileio.statis_equal(vals[k], want[k], 1e-5)]
    assert not bad, bad
