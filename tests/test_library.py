"""CPU checks of the C ABI library and the host-side Python: the library loads,
exports every function include/dlpx.h declares, its structs have the layout the
binding assumes, and the system generator / file formats behave.  No compute
calls (no GPU here)."""
import ctypes as C
import os
import re
import subprocess
import tempfile

import numpy as np
import pytest

from dlpx import fileio, lib, systems

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "dlpx.h")


def header_functions():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char \*)\s*(dlpx_\w+)\s*\(", text, re.M)))


def test_library_exports_every_declared_symbol():
    names = header_functions()
    assert len(names) >= 18
    so = lib.load()
    for n in names:
        assert hasattr(so, n), n
    assert sorted(lib.EXPORTS) == names
    out = subprocess.run(["nm", "-D", "--defined-only", lib.LIB], capture_output=True, text=True).stdout
    exported = set(re.findall(r"\bT (dlpx_\w+)", out))
    assert set(names) <= exported


def test_struct_layout_matches_header():
    src = r"""
#include <stdio.h>
#include <stddef.h>
#include "dlpx.h"
int main(void) {
  printf("%zu %zu %zu %zu %zu %zu\n", sizeof(dlpx_params), offsetof(dlpx_params, kvec),
         offsetof(dlpx_params, epsilon), sizeof(dlpx_atoms), offsetof(dlpx_atoms, ltype),
         sizeof(dlpx_two_body_out));
  return 0;
}
"""
    with tempfile.TemporaryDirectory() as d:
        c = os.path.join(d, "t.c")
        open(c, "w").write(src)
        exe = os.path.join(d, "t")
        subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), c, "-o", exe], check=True)
        got = [int(v) for v in subprocess.run([exe], capture_output=True, text=True).stdout.split()]
    want = [C.sizeof(lib.Params), lib.Params.kvec.offset, lib.Params.epsilon.offset, C.sizeof(lib.Atoms),
            lib.Atoms.ltype.offset, C.sizeof(lib.TwoBodyOut)]
    assert got == want


def test_nacl_generator():
    s = systems.nacl(ncells=4, kvec=(32, 32, 32))
    assert s.natms == 512 and abs(s.q.sum()) < 1e-12
    L = s.cell[0]
    assert abs(L - 4 * 5.64 * 1.08) < 1e-12
    assert np.all(s.pos >= -L / 2) and np.all(s.pos < L / 2)
    assert np.all(np.diff(s.types) >= 0)  # Na then Cl
    # centre-of-mass momentum removed
    assert np.abs((s.mass[:, None] * s.vel).sum(0)).max() < 1e-9
    big = systems.nacl(ncells=20)
    assert big.natms == 64000 and big.kvec == (125, 125, 125)
    assert systems.smooth235_at_least(121) == 125 and systems.smooth235_at_least(361) == 375


def test_argon_density():
    s = systems.argon()
    L = s.cell[0]
    rho = s.natms * systems.AR_MASS * 1.66053906660 / L ** 3
    assert s.natms == 4000 and abs(rho - 1.374) < 1e-12 and abs(L - 57.8015) < 1e-3


def test_config_round_trip():
    s = systems.nacl(ncells=3, kvec=(20, 20, 20), cutoff=8.5)
    with tempfile.TemporaryDirectory() as d:
        fileio.write_inputs(d, s)
        txt = open(os.path.join(d, "CONFIG")).read().splitlines()
        assert all(len(ln) <= 72 for ln in txt)
        pos = np.array([[float(v) for v in txt[5 + 3 * i + 1].split()] for i in range(s.natms)])
        assert np.array_equal(pos, s.pos)  # shortest round-trip repr: bit-exact
        field = open(os.path.join(d, "FIELD")).read()
        assert "units internal" in field and field.count("bhm") == 3
        ctrl = open(os.path.join(d, "CONTROL")).read()
        assert "spme_kvec [20 20 20]" in ctrl and "restart noscale" in ctrl


def test_revive_reader():
    with tempfile.TemporaryDirectory() as d:
        p = os.path.join(d, "REVIVE")
        recs = [np.arange(3.0), np.arange(8.0), np.zeros(10), np.arange(55.0) * 1.5] + \
               [np.zeros(5)] * 8 + [np.arange(9.0) + 0.25]
        with open(p, "wb") as f:
            for r in recs:
                b = r.astype("<f8").tobytes()
                f.write(np.int32(len(b)).tobytes() + b + np.int32(len(b)).tobytes())
        rv = fileio.read_revive(p)
        assert rv["engcpe"] == 5 * 1.5 and rv["vircpe"] == 14 * 1.5
        assert np.array_equal(rv["stress"], np.arange(9.0) + 0.25)


def test_missing_library_fails_loudly(monkeypatch):
    monkeypatch.setattr(lib, "_lib", None)
    monkeypatch.setattr(lib, "LIB", "/nonexistent/libdlpx.so")
    with pytest.raises(ImportError):
        lib.load()


def test_fortran_binding_builds_and_links():
    exe = os.path.join(ROOT, "dl-poly_amd", "fortran", "dlpx_host_demo")
    assert os.path.exists(exe), "built by __graft_entry__.build()"
    out = subprocess.run(["ldd", exe], capture_output=True, text=True).stdout
    assert "libdlpx.so" in out and "not found" not in out
