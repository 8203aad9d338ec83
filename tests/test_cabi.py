"""The C-ABI library loads and exports every symbol include/h3dmg.h declares
(no compute calls: runs without a GPU)."""
import ctypes
import os
import re

from hydro3d_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "h3dmg.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(h3dmg_[a-z_0-9]+)\s*\(", text)))


def test_header_declares_the_boundary():
    names = declared_functions()
    for must in ("h3dmg_create", "h3dmg_set_rhs", "h3dmg_vcycle", "h3dmg_solve", "h3dmg_get_pp",
                 "h3dmg_destroy", "h3dmg_last_error"):
        assert must in names


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(_lib.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_python_binding_covers_every_symbol():
    assert set(declared_functions()) == set(_lib.SIGNATURES)


def test_struct_layouts_match_header():
    # h3dmg_block: 3 ints, 3 doubles, int, 6 ints
    assert ctypes.sizeof(_lib.H3dBlock) == 4 * 3 + 4 + 8 * 3 + 4 + 4 * 6 + 4
    # 5 ints + pad, 2 doubles, 4 ints
    assert ctypes.sizeof(_lib.H3dParams) == 24 + 16 + 16


def test_default_params_without_device():
    lib = _lib.load()
    p = _lib.H3dParams()
    lib.h3dmg_default_params(ctypes.byref(p))
    assert (p.smoother, p.nu1, p.nu2, p.maxcy) == (3, 2, 2, 60)
    assert p.tol == 1e-8 and p.precision == _lib.FP64


def test_create_rejects_bad_levels_before_touching_the_device():
    lib = _lib.load()
    blk = (_lib.H3dBlock * 1)()
    blk[0].ni, blk[0].nj, blk[0].nk = 12, 8, 8
    blk[0].dx = blk[0].dy = blk[0].dz = 1.0
    blk[0].nlevels = 4
    h = ctypes.c_void_p()
    st = lib.h3dmg_create(blk, 1, None, ctypes.byref(h))
    # no device here: either the device probe or the level rule refuses
    assert st in (_lib.ERR_LEVELS, _lib.ERR_DEVICE)
    assert not h.value
