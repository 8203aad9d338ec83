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
    # h3dmg_block: 3 ints (+pad), 3 doubles, int, 6 ints, 6 ints (+pad)
    assert ctypes.sizeof(_lib.H3dBlock) == 4 * 3 + 4 + 8 * 3 + 4 + 4 * 6 + 4 * 6 + 4
    # 5 ints + pad, 2 doubles, 4 ints
    assert ctypes.sizeof(_lib.H3dParams) == 24 + 16 + 16


def test_default_params_without_device():
    lib = _lib.load()
    p = _lib.H3dParams()
    lib.h3dmg_default_params(ctypes.byref(p))
    assert (p.smoother, p.nu1, p.nu2, p.maxcy) == (3, 2, 2, 60)
    assert p.tol == 1e-8 and p.precision == _lib.FP64


def _one_block(ni, nj, nk, nlev, bc=(0, 0, 0, 0, 0, 0)):
    blk = (_lib.H3dBlock * 1)()
    blk[0].ni, blk[0].nj, blk[0].nk = ni, nj, nk
    blk[0].dx = blk[0].dy = blk[0].dz = 1.0
    blk[0].nlevels = nlev
    for f in range(6):
        blk[0].bc[f] = bc[f]
    return blk


def test_create_rejects_bad_levels_before_touching_the_device():
    """localparameters.for:202-208: every axis divisible by 2^(ngrid-1), or
    the run stops.  The rule is checked on the host before any device call,
    so the refusal is ERR_LEVELS here, without a GPU."""
    lib = _lib.load()
    h = ctypes.c_void_p()
    st = lib.h3dmg_create(_one_block(12, 8, 8, 4), 1, None, ctypes.byref(h))
    assert st == _lib.ERR_LEVELS, lib.h3dmg_last_error()
    assert b"not divisible by 2^(nlevels-1) = 8" in lib.h3dmg_last_error()
    assert not h.value
    # the last axis alone breaks the rule
    st = lib.h3dmg_create(_one_block(16, 16, 20, 4), 1, None, ctypes.byref(h))
    assert st == _lib.ERR_LEVELS and not h.value


def test_create_rejects_bad_blocks_before_touching_the_device():
    lib = _lib.load()
    h = ctypes.c_void_p()
    # a periodic axis needs both faces periodic
    st = lib.h3dmg_create(_one_block(8, 8, 8, 2, bc=(1, 0, 0, 0, 0, 0)), 1, None,
                          ctypes.byref(h))
    assert st == _lib.ERR_ARG and b"periodic on one face" in lib.h3dmg_last_error()
    st = lib.h3dmg_create(_one_block(8, 8, 8, 0), 1, None, ctypes.byref(h))
    assert st == _lib.ERR_ARG and b"nlevels" in lib.h3dmg_last_error()
    # mixed precision needs a coarser level to run in FP32
    p = _lib.H3dParams()
    lib.h3dmg_default_params(ctypes.byref(p))
    p.precision = _lib.MIXED
    st = lib.h3dmg_create(_one_block(8, 8, 8, 1), 1, ctypes.byref(p), ctypes.byref(h))
    assert st == _lib.ERR_ARG and b"mixed precision" in lib.h3dmg_last_error()
    assert not h.value
