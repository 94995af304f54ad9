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


# Translation units of the library: each kernel family compiles on its own,
# in parallel, and the objects are linked into one shared library.
SOURCES = ["hc_api.hip", "hc_xfer.hip", "hc_rhs.hip", "hc_rhs_t.hip", "hc_columns.hip", "hc_spec.hip",
           "hc_spec_col.hip", "hc_z6.hip", "hc_adj.hip", "hc_screen.hip", "hc_stages.hip", "hc_stages_sp.hip"]
CSRC = os.path.join(PKG, "csrc")
OBJ = os.path.join(PKG, "_obj")


def _stale(src, obj, deps):
    if not os.path.exists(obj):
        return True
    t = os.path.getmtime(obj)
    return any(os.path.getmtime(d) > t for d in [src, *deps])


def build_hip(verbose=False, incremental=False, jobs=None):
    """Compile every unit (or, with incremental=True, those whose source or any
    header changed since their object was built) and link the library."""
    os.makedirs(OBJ, exist_ok=True)
    deps = [os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".h")]
    deps.append(os.path.join(ROOT, "include", "hercules_amd.h"))
    objs, todo = [], []
    for f in SOURCES:
        src = os.path.join(CSRC, f)
        obj = os.path.join(OBJ, f.replace(".hip", ".o"))
        objs.append(obj)
        if not incremental or _stale(src, obj, deps):
            todo.append((src, obj))
    flags = [x for x in HIP_FLAGS if x != "-shared"]

    def one(so):
        src, obj = so
        tmp = obj + ".tmp.o"
        _run([HIPCC, *flags, "-c", "-o", tmp, src])
        os.replace(tmp, obj)
        return src

    from concurrent.futures import ThreadPoolExecutor
    with ThreadPoolExecutor(max_workers=jobs or min(len(todo) or 1, os.cpu_count() or 4)) as ex:
        for src in ex.map(one, todo):
            if verbose:
                print("compiled", os.path.basename(src))
    tmp = LIB + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs])
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    return LIB


FC = os.environ.get("HC_FC", shutil.which("amdflang") or "/opt/rocm/bin/amdflang")
FDIR = os.path.join(PKG, "fortran")
FEXE = os.path.join(FDIR, "hercules_hip")
STAGE_SRC = os.path.join(ROOT, "tests", "fortran", "stage_loop.f90")
STAGE_EXE = os.path.join(ROOT, "tests", "fortran", "stage_loop")
STAGE_SP_SRC = os.path.join(ROOT, "tests", "fortran", "stage_loop_sp.f90")
STAGE_SP_EXE = os.path.join(ROOT, "tests", "fortran", "stage_loop_sp")


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
    # the tests' host-loop callers of the per-stage subroutines (FD and spectral)
    for src, exe in ((STAGE_SRC, STAGE_EXE), (STAGE_SP_SRC, STAGE_SP_EXE)):
        if os.path.exists(src):
            _run([FC, "-O2", src, obj, "-module-dir", out, "-I", out, "-o", exe,
                  "-L" + PKG, "-lhercules_amd", "-Wl,-rpath," + PKG])
            if verbose:
                print("built", exe)
    return FEXE


def build_all(verbose=False):
    build_hip(verbose)
    build_fortran(verbose)


if __name__ == "__main__":
    import sys
    if "--incremental" in sys.argv:
        build_hip(verbose=True, incremental=True)
    else:
        build_all(verbose=True)
