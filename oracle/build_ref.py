#!/usr/bin/env python3
"""Compile the unmodified DL_POLY 5.1 reference (serial build) into oracle/_ref/.

Test infrastructure only.  The reference is Fortran 2008; this image ships
AMD flang (``/opt/rocm/bin/amdflang``).  The reference's own CMake build is not
used: this script reads the source list the reference declares
(``source/CMakeLists.txt`` ``dlpoly_SRCS`` plus ``source/SERIAL/mpi_api.F90``
and the ``dl_poly.F90`` main program), orders the files by their ``Use``
statements and compiles them where they lie under /root/reference.  Every
output (objects, ``.mod`` files, ``DLPOLY.Z``) goes to ``oracle/_ref/``.

Two compiler-side settings are needed, both non-semantic and neither touching
the sources:

* ``vdw.F90`` declares two MDF long-range-correction integrals ``Pure`` while
  they assign a polymorphic allocatable (``vdw.F90:376,449``), which flang
  rejects.  That one file is compiled with ``-DPure=`` so the preprocessor drops
  the ``Pure`` attribute; purity never changes a computed value.
* ``electrostatic.F90`` passes internal procedures as actual arguments
  (``electrostatic.F90:107-108``); their trampolines need an executable stack,
  hence ``-Wl,-z,execstack`` at link time.

Floating point is compiled ``-O2`` without fast-math so the oracle's
arithmetic is IEEE-faithful.

Usage: python oracle/build_ref.py [--jobs N]
"""
import argparse
import concurrent.futures as cf
import os
import re
import subprocess
import sys

REF = "/root/reference"
SRC = os.path.join(REF, "source")
HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, "_ref")
OBJ = os.path.join(OUT, "obj")
FC = os.environ.get("DLPX_FC", "/opt/rocm/bin/amdflang")
FFLAGS = ["-O2", "-cpp", "-DSERIAL", "-DUNIX", "-DCHRONO", "-DVERSION='5.1.0'",
          "-DVERSION_DATE='oracle'", "-DHOST='oracle'", "-I", os.path.join(SRC, "SERIAL")]
PER_FILE = {"vdw.F90": ["-DPure="]}


def source_list():
    """Files named by the reference's dlpoly_SRCS list (source/CMakeLists.txt:3-123)."""
    text = open(os.path.join(SRC, "CMakeLists.txt")).read()
    block = text.split("set(dlpoly_SRCS", 1)[1].split(")", 1)[0]
    files = []
    for line in block.splitlines():
        line = line.strip()
        if not line or line.startswith("#"):
            continue
        files.append(line)
    files.append("SERIAL/mpi_api.F90")
    return files


MOD_RE = re.compile(r"^\s*module\s+(\w+)\s*$", re.I)
USE_RE = re.compile(r"^\s*use\s*(?:,\s*\w+\s*)?(?:::)?\s*(\w+)", re.I)


def scan(path):
    mods, uses = set(), set()
    with open(path, errors="replace") as f:
        for line in f:
            m = MOD_RE.match(line)
            if m and m.group(1).lower() != "procedure":
                mods.add(m.group(1).lower())
            u = USE_RE.match(line)
            if u:
                uses.add(u.group(1).lower())
    return mods, uses


def levels(files):
    info = {f: scan(os.path.join(SRC, f)) for f in files}
    provider = {}
    for f, (mods, _) in info.items():
        for m in mods:
            provider[m] = f
    deps = {f: {provider[u] for u in uses if u in provider and provider[u] != f}
            for f, (_, uses) in info.items()}
    done, out = set(), []
    while len(done) < len(files):
        ready = sorted(f for f in files if f not in done and deps[f] <= done)
        if not ready:
            raise RuntimeError("cyclic module dependencies: %s" % sorted(set(files) - done))
        out.append(ready)
        done |= set(ready)
    return out


def compile_one(f):
    src = os.path.join(SRC, f)
    obj = os.path.join(OBJ, os.path.basename(f).replace(".F90", ".o"))
    if os.path.exists(obj) and os.path.getmtime(obj) > os.path.getmtime(src):
        return obj
    cmd = [FC] + FFLAGS + PER_FILE.get(os.path.basename(f), []) + \
          ["-module-dir", OBJ, "-I", OBJ, "-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("compile failed: %s\n%s\n%s" % (" ".join(cmd), r.stdout, r.stderr[-4000:]))
    return obj


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=8)
    args = ap.parse_args()
    if not os.path.isdir(SRC):
        sys.exit("reference sources not found at %s" % SRC)
    os.makedirs(OBJ, exist_ok=True)
    files = source_list()
    objs = []
    for lvl in levels(files):
        with cf.ThreadPoolExecutor(args.jobs) as ex:
            objs += list(ex.map(compile_one, lvl))
    main_obj = compile_one("dl_poly.F90")
    exe = os.path.join(OUT, "DLPOLY.Z")
    cmd = [FC, "-o", exe, main_obj] + objs + ["-Wl,-z,execstack"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n%s" % r.stderr[-4000:])
    print(exe)


if __name__ == "__main__":
    main()
