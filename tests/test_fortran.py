"""Fortran host binding (hydro3d_amd/fortran/h3dmg_mod.f90).

CPU: the binding and the drop-in newsolv_mg_h3d compile, and the demo
program runs through the C-ABI far enough to get a clean error status when
no GPU is present.  GPU: the demo solves a discrete eigenmode (closed-form
solution) through the Fortran interface: red-black in FP64 and mixed
precision, lexicographic GS and line TDMA in FP64.
"""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FDIR = os.path.join(ROOT, "hydro3d_amd", "fortran")
DEMO = os.path.join(FDIR, "h3dmg_fortran_demo")
SHIM_OBJ = os.path.join(FDIR, "_build", "newsolv_mg_h3d.o")

needs_demo = pytest.mark.skipif(not os.path.exists(DEMO),
                                reason="Fortran demo not built (flang absent at build time)")


@needs_demo
def test_fortran_demo_reports_missing_device_cleanly():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present: the GPU test runs the demo")
    r = subprocess.run([DEMO], capture_output=True, text=True, timeout=60)
    assert r.returncode != 0
    assert "h3dmg_create failed: status  3" in r.stdout, r.stdout


def test_dropin_shim_compiled_when_reference_modules_exist():
    mods = os.path.join(ROOT, "oracle", "_ref", "obj", "multidata.mod")
    if not os.path.exists(mods):
        pytest.skip("reference modules not built here")
    assert os.path.exists(SHIM_OBJ), "newsolv_mg_h3d.f90 was not compiled by build()"


@pytest.mark.gpu
@needs_demo
def test_fortran_demo_solves_eigenmode_on_gpu():
    print("case: Fortran demo 32^3 PPN eigenmode, rbgs fp64 + mixed, gs, tdma", flush=True)
    r = subprocess.run([DEMO], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "h3dmg fortran demo: OK" in r.stdout
