"""Build recipes of the native parts (run on the CPU host; hipcc
cross-compiles gfx950 code without a GPU).

    python -m hydro3d_amd.build          # libh3dmg.so (+ Fortran binding)

Outputs are written in-tree next to the package so that they travel with the
source tree to the machine that runs them.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
LIB = os.path.join(PKG, "libh3dmg.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLANG = os.environ.get("FLANG", "/opt/rocm/lib/llvm/bin/flang")

HIP_SOURCES = ["h3dmg_kernels.hip", "h3dmg_runtime.hip", "h3dmg_layout.cpp", "h3dmg_tecbin.cpp"]

# -ffp-contract=off: the FP64 path must execute the reference's products and
# sums as written (no FMA contraction) to stay bitwise equal to it.
# -packed-fp32-ops off: no v_pk_* FP32 arithmetic is generated.
HIP_FLAGS = [
    "-O3", "-std=c++17", "--offload-arch=gfx950", "-fPIC",
    "-ffp-contract=off", "-fno-fast-math",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-result",
    "-mllvm", "-amdgpu-kernarg-preload-count=16",
]


def _run(cmd, cwd=None):
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True, cwd=cwd)


def build_lib(verbose=True) -> str:
    objdir = os.path.join(PKG, "_build")
    os.makedirs(objdir, exist_ok=True)
    objs = []
    for src in HIP_SOURCES:
        obj = os.path.join(objdir, os.path.splitext(src)[0] + ".o")
        _run([HIPCC, *HIP_FLAGS, "-I", os.path.join(ROOT, "include"), "-c",
              os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    tmp = LIB + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs])
    os.replace(tmp, LIB)
    return LIB


PROBES = ["probe_grid_barrier"]


def build_probes() -> list:
    """Measurement programs beside the library (executables in the package
    directory): probe_grid_barrier times a device-scope grid barrier against
    a kernel boundary in a captured graph."""
    out = []
    for name in PROBES:
        exe = os.path.join(PKG, name)
        _run([HIPCC, *HIP_FLAGS, "-o", exe, os.path.join(CSRC, name + ".hip")])
        out.append(exe)
    return out


def build_fortran() -> str | None:
    """ISO_C_BINDING module and the Fortran test program (needs flang)."""
    fdir = os.path.join(PKG, "fortran")
    if not os.path.exists(FLANG) or not os.path.isdir(fdir):
        return None
    src = os.path.join(fdir, "h3dmg_mod.f90")
    if not os.path.exists(src):
        return None
    out = os.path.join(fdir, "_build")
    os.makedirs(out, exist_ok=True)
    _run([FLANG, "-O2", "-module-dir", out, "-c", src, "-o", os.path.join(out, "h3dmg_mod.o")])
    # the drop-in newsolv_mg_h3d needs the reference's modules (vars, mpi,
    # multidata): compile-checked where the reference objects were built
    ref_mods = os.path.join(os.path.dirname(PKG), "oracle", "_ref", "obj")
    if os.path.exists(os.path.join(ref_mods, "multidata.mod")):
        for name in ("newsolv_mg_h3d", "newsolv_mg"):
            _run([FLANG, "-O2", "-ffp-contract=off", "-fdefault-real-8", "-fdefault-double-8",
                  "-I", ref_mods, "-I", out, "-I/opt/conda/include", "-c",
                  os.path.join(fdir, name + ".f90"), "-o", os.path.join(out, name + ".o")])
    prog = os.path.join(fdir, "h3dmg_fortran_demo.f90")
    if os.path.exists(prog):
        exe = os.path.join(fdir, "h3dmg_fortran_demo")
        _run([FLANG, "-O2", "-I", out, "-o", exe, prog, os.path.join(out, "h3dmg_mod.o"),
              "-L", PKG, "-Wl,-rpath,$ORIGIN/..", "-lh3dmg"])
        return exe
    return None


def main(argv=None):
    build_lib()
    build_probes()
    build_fortran()


if __name__ == "__main__":
    main(sys.argv[1:])
