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


def test_block_bytes_is_host_only_and_counts_the_region():
    """h3dmg_block_bytes (a pooled block's region) needs no device: per
    level six colour-split arrays (phi twice, rhs; red and black halves of
    H = (ni+1)//2 + 1 rounded up to even entries per row), FP32 below the
    coarsest level in mixed precision; four FP64 finest arrays (p, f); the
    exact coarse solve's eigen tables; each array 256-byte aligned."""
    from hydro3d_amd.solver import block_bytes

    def al(x):
        return (x + 255) // 256 * 256

    def expect(n, nlev, mixed, exact):
        tot = 0
        for g in range(1, nlev + 1):
            ni, nj, nk = (x >> (g - 1) for x in n)
            half = (((ni + 1) // 2 + 1) & ~1) * nj * nk * (4 if mixed and g < nlev else 8)
            tot += 6 * al(half)
        ni, nj, nk = n
        tot += 4 * al((((ni + 1) // 2 + 1) & ~1) * nj * nk * 8)
        if exact:
            c = [x >> (nlev - 1) for x in n]
            tot += sum(al(x * x * 8) + al(x * 8) for x in c) + 2 * al(c[0] * c[1] * c[2] * 8)
        return tot

    for n, nlev in (((64, 48, 32), 4), ((96, 64, 64), 5), ((256, 256, 256), 6), ((48, 80, 16), 3)):
        spec = dict(n=n, h=(0.01,) * 3, bc=(0,) * 6, nlev=nlev)
        assert block_bytes(spec, precision="mixed", coarse="exact") == expect(n, nlev, True, True)
        assert block_bytes(spec, precision="fp64", coarse="smooth") == expect(n, nlev, False, False)
    lib = _lib.load()
    p = _lib.H3dParams()
    lib.h3dmg_default_params(ctypes.byref(p))
    assert lib.h3dmg_block_bytes(_one_block(12, 8, 8, 4), ctypes.byref(p)) == 0
    assert b"not divisible" in lib.h3dmg_last_error()
