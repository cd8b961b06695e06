#!/usr/bin/env python3
"""Link the compiled DL_POLY reference (oracle/build_ref.py's objects) with the
dump hook of tests/golden/ref_dump.F90 and ref_wrap.c into
oracle/_ref/DLPOLY_DUMP.Z.

Test infrastructure only, run in the build container: the binary is the
reference itself, with two_body_forces wrapped at link time (GNU ld --wrap) so
that after its first call the forces, energies, stress and vdW tables are
written at full precision (tests/golden/make_fixtures.py reads them).  No
reference source is copied or edited.

Usage: python tests/golden/build_ref_dump.py
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REF = os.path.join(ROOT, "oracle", "_ref")
OBJ = os.path.join(REF, "obj")
FC = os.environ.get("DLPX_FC", "/opt/rocm/bin/amdflang")
SYM = "_QMtwo_bodyPtwo_body_forces"


def run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("failed: %s\n%s\n%s" % (" ".join(cmd), r.stdout, r.stderr[-4000:]))


def main():
    if not os.path.exists(os.path.join(OBJ, "two_body.o")):
        sys.exit("build the reference first: python oracle/build_ref.py")
    dump_dir = os.path.join(REF, "dump")
    os.makedirs(dump_dir, exist_ok=True)
    dump_o, wrap_o = os.path.join(dump_dir, "ref_dump.o"), os.path.join(dump_dir, "ref_wrap.o")
    run([FC, "-O2", "-I", OBJ, "-module-dir", dump_dir, "-c", os.path.join(HERE, "ref_dump.F90"), "-o", dump_o])
    run(["gcc", "-O2", "-c", os.path.join(HERE, "ref_wrap.c"), "-o", wrap_o])
    objs = sorted(os.path.join(OBJ, f) for f in os.listdir(OBJ) if f.endswith(".o"))
    exe = os.path.join(REF, "DLPOLY_DUMP.Z")
    run([FC, "-o", exe] + objs + [dump_o, wrap_o, "-Wl,--wrap=" + SYM, "-Wl,-z,execstack"])
    print(exe)


if __name__ == "__main__":
    main()
