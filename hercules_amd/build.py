"""Build recipe for the native parts of hercules_amd.

* ``libhercules_amd.so`` — the HIP kernels and the C ABI of
  ``include/hercules_amd.h``, compiled for gfx950 with hipcc.
Outputs are written next to the sources (in-tree) so they travel with the
repository snapshot.
"""
from __future__ import annotations

import os
import shutil
import subprocess

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
LIB = os.path.join(PKG, "libhercules_amd.so")

HIPCC = os.environ.get("HIPCC", shutil.which("hipcc") or "/opt/rocm/bin/hipcc")
HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    # packed FP32 instructions are kept out of every kernel
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def _run(cmd, cwd=None):
    r = subprocess.run(cmd, cwd=cwd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"build failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return r


def build_hip(verbose=False):
    src = os.path.join(PKG, "csrc", "hc_core.hip")
    tmp = LIB + ".tmp"
    _run([HIPCC, *HIP_FLAGS, "-o", tmp, src])
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    return LIB


FC = os.environ.get("HC_FC", shutil.which("amdflang") or "/opt/rocm/bin/amdflang")
FDIR = os.path.join(PKG, "fortran")
FEXE = os.path.join(FDIR, "hercules_hip")


def build_fortran(verbose=False):
    """The Fortran host: ISO_C_BINDING module + driver, linked to the HIP library."""
    out = os.path.join(FDIR, "_mod")
    os.makedirs(out, exist_ok=True)
    mod = os.path.join(FDIR, "module_hercules_hip.f90")
    main = os.path.join(FDIR, "hercules_hip_main.f90")
    obj = os.path.join(out, "module_hercules_hip.o")
    _run([FC, "-O2", "-c", mod, "-module-dir", out, "-o", obj])
    _run([FC, "-O2", main, obj, "-module-dir", out, "-I", out, "-o", FEXE,
          "-L" + PKG, "-lhercules_amd", "-Wl,-rpath," + PKG, "-Wl,-rpath,$ORIGIN/.."])
    if verbose:
        print("built", FEXE)
    return FEXE


def build_all(verbose=False):
    build_hip(verbose)
    build_fortran(verbose)


if __name__ == "__main__":
    build_all(verbose=True)
