"""Compile the HIP sources of the force path into ``libdlpx.so`` (gfx950).

The library is built in-tree next to this file so it travels with the
repository; nothing is installed.  Usage: ``python -m dlpx.build`` or
``dlpx.build.build()``.
"""
import concurrent.futures as cf
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.dirname(HERE)
CSRC = os.path.join(PKG, "csrc")
ROOT = os.path.dirname(PKG)
INCLUDE = os.path.join(ROOT, "include")
OBJ = os.path.join(PKG, "_build")
LIB = os.path.join(HERE, "libdlpx.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("DLPX_ARCH", "gfx950")
FLAGS = ["-O3", "-fPIC", "-std=c++17", "--offload-arch=" + ARCH, "-I", INCLUDE, "-I", CSRC,
         "-munsafe-fp-atomics", "-Wall", "-Wno-unused-function",
         # no packed FP32 instructions (v_pk_*_f32) in generated code
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]
SOURCES = ["tables.cpp", "neighbours.hip", "pair.hip", "spme.hip", "fft.hip", "fft_lines.hip", "fft_planes.hip", "nve.hip", "api.cpp", "batch.cpp"]


def _compile(src):
    out = os.path.join(OBJ, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC] + FLAGS + ["-x", "hip", "-c", os.path.join(CSRC, src), "-o", out]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed for %s:\n%s" % (src, r.stderr[-6000:]))
    return out


FC = os.environ.get("DLPX_FC", "/opt/rocm/bin/amdflang")
FORTRAN = os.path.join(PKG, "fortran")


def build_fortran(verbose=False):
    """The ISO_C_BINDING module and the Fortran host driver, linked to libdlpx.so."""
    if not os.path.exists(FC):
        if verbose:
            print("no Fortran compiler at %s: Fortran host driver not built" % FC)
        return None
    mod = os.path.join(OBJ, "fortran")
    os.makedirs(mod, exist_ok=True)
    obj = os.path.join(mod, "dlpx_binding.o")
    exe = os.path.join(FORTRAN, "dlpx_host_demo")
    for cmd in ([FC, "-O2", "-c", os.path.join(FORTRAN, "dlpx_binding.F90"), "-module-dir", mod, "-o", obj],
                [FC, "-O2", "-I", mod, os.path.join(FORTRAN, "dlpx_host_demo.F90"), obj, "-L", HERE, "-ldlpx",
                 "-Wl,-rpath,$ORIGIN/../dlpx", "-o", exe]):
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError("Fortran build failed:\n%s\n%s" % (" ".join(cmd), r.stderr[-4000:]))
    if verbose:
        print(exe)
    return exe


def build(jobs=8, verbose=False):
    os.makedirs(OBJ, exist_ok=True)
    with cf.ThreadPoolExecutor(jobs) as ex:
        objs = list(ex.map(_compile, SOURCES))
    cmd = [HIPCC, "-shared", "-fPIC", "--offload-arch=" + ARCH, "-o", LIB + ".tmp"] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n%s" % r.stderr[-4000:])
    os.replace(LIB + ".tmp", LIB)
    if verbose:
        print(LIB)
    build_fortran(verbose)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
