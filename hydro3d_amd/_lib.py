"""ctypes binding of libh3dmg.so (the C-ABI declared in include/h3dmg.h).

The library is loaded from the package directory only; if it is missing the
import fails loudly -- there is no CPU fallback behind this module.
"""
from __future__ import annotations

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libh3dmg.so")

OK = 0
ERR_ARG, ERR_LEVELS, ERR_DEVICE, ERR_NOMEM = 1, 2, 3, 4
ERR_NOT_CONVERGED, ERR_DIVERGED, ERR_UNSUPPORTED, ERR_BUSY = 5, 6, 7, 8

BC_WALL, BC_PERIODIC, BC_NEIGHBOUR = 0, 1, 2
SMOOTH_GS, SMOOTH_TDMA, SMOOTH_RBGS = 1, 2, 3
CRIT_REL_L2, CRIT_REF_RMAX = 0, 1
FP64, MIXED = 0, 1
COARSE_SMOOTH, COARSE_EXACT = 0, 1
ARR_PHI, ARR_RHS = 0, 1


class H3dBlock(ctypes.Structure):
    _fields_ = [("ni", ctypes.c_int), ("nj", ctypes.c_int), ("nk", ctypes.c_int),
                ("dx", ctypes.c_double), ("dy", ctypes.c_double), ("dz", ctypes.c_double),
                ("nlevels", ctypes.c_int), ("bc", ctypes.c_int * 6), ("nbr", ctypes.c_int * 6)]


TECBIN_FIELDS = ("p", "pm", "ppm", "vis", "vism", "u", "um", "uum", "v", "vm", "vvm", "w", "wm",
                 "wwm", "uvm", "uwm", "vwm", "ksgs", "eps", "epsm")


class H3dTecbinHeader(ctypes.Structure):
    _fields_ = [("tt", ctypes.c_int * 3), ("pl", ctypes.c_int), ("ind", ctypes.c_int * 3)]


class H3dParams(ctypes.Structure):
    _fields_ = [("smoother", ctypes.c_int), ("nu1", ctypes.c_int), ("nu2", ctypes.c_int),
                ("maxcy", ctypes.c_int), ("criterion", ctypes.c_int), ("tol", ctypes.c_double),
                ("dt", ctypes.c_double), ("precision", ctypes.c_int), ("coarse", ctypes.c_int),
                ("min_cycles", ctypes.c_int), ("use_graph", ctypes.c_int)]


class H3dBlockState(ctypes.Structure):
    _fields_ = [("running", ctypes.c_int), ("cycles", ctypes.c_int), ("status", ctypes.c_int),
                ("enabled", ctypes.c_int), ("fnorm", ctypes.c_double), ("rel", ctypes.c_double),
                ("tol", ctypes.c_double)]


# every symbol the header declares: name -> (restype, argtypes)
_P = ctypes.POINTER
_vp = ctypes.c_void_p
_dp = _P(ctypes.c_double)
_ip = _P(ctypes.c_int)
_i = ctypes.c_int
_d = ctypes.c_double
SIGNATURES = {
    "h3dmg_default_params": (None, [_P(H3dParams)]),
    "h3dmg_create": (_i, [_P(H3dBlock), _i, _P(H3dParams), _P(_vp)]),
    "h3dmg_destroy": (None, [_vp]),
    "h3dmg_last_error": (ctypes.c_char_p, []),
    "h3dmg_device_bytes": (ctypes.c_ulonglong, [_vp]),
    "h3dmg_set_rhs": (_i, [_vp, _i, _dp, _i, _i, _i]),
    "h3dmg_set_rhs_device": (_i, [_vp, _i, _vp]),
    "h3dmg_set_p": (_i, [_vp, _i, _dp, _i, _i, _i]),
    "h3dmg_get_p": (_i, [_vp, _i, _dp, _i, _i, _i]),
    "h3dmg_set_p_device": (_i, [_vp, _i, _vp]),
    "h3dmg_get_p_device": (_i, [_vp, _i, _vp]),
    "h3dmg_vcycle": (_i, [_vp]),
    "h3dmg_get_pp": (_i, [_vp, _i, _dp, _i, _i, _i]),
    "h3dmg_solve": (_i, [_vp, _dp, _i, _ip, _dp]),
    "h3dmg_block_stats": (_i, [_vp, _ip, _dp]),
    "h3dmg_remove_mean": (_i, [_vp]),
    "h3dmg_set_stream": (_i, [_vp, _vp]),
    "h3dmg_set_march": (_i, [_vp, _i, _i]),
    "h3dmg_get_march": (_i, [_vp, _i]),
    "h3dmg_set_lanes": (_i, [_vp, _i, _i]),
    "h3dmg_get_lanes": (_i, [_vp, _i]),
    "h3dmg_set_fused": (_i, [_vp, ctypes.c_longlong]),
    "h3dmg_get_fused": (ctypes.c_longlong, [_vp]),
    "h3dmg_set_tail": (_i, [_vp, _i]),
    "h3dmg_set_block_tol": (_i, [_vp, _i, _d]),
    "h3dmg_set_block_enabled": (_i, [_vp, _i, _i]),
    "h3dmg_pool_create": (_i, [ctypes.c_ulonglong, _i, _P(_vp)]),
    "h3dmg_pool_destroy": (None, [_vp]),
    "h3dmg_pool_alloc": (_i, [_vp, ctypes.c_ulonglong, _P(ctypes.c_ulonglong)]),
    "h3dmg_pool_free": (_i, [_vp, ctypes.c_ulonglong]),
    "h3dmg_pool_free_bytes": (ctypes.c_ulonglong, [_vp]),
    "h3dmg_pool_largest_free": (ctypes.c_ulonglong, [_vp]),
    "h3dmg_block_bytes": (ctypes.c_ulonglong, [_P(H3dBlock), _P(H3dParams)]),
    "h3dmg_create_pooled": (_i, [_vp, _P(H3dBlock), _i, _P(H3dParams), _P(ctypes.c_ulonglong),
                                 _ip, _vp, _P(_vp)]),
    "h3dmg_get_block_state": (_i, [_vp, _i, _P(H3dBlockState)]),
    "h3dmg_set_block_state": (_i, [_vp, _i, _P(H3dBlockState)]),
    "h3dmg_solve_begin": (_i, [_vp, _ip, _i]),
    "h3dmg_solve_step": (_i, [_vp, _ip]),
    "h3dmg_project_device": (_i, [_vp, _i, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _d, _dp]),
    "h3dmg_project_device_noalloc": (_i, [_vp, _i, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _d, _vp]),
    "h3dmg_layout_read": (_i, [ctypes.c_char_p, ctypes.c_char_p, _P(_vp)]),
    "h3dmg_layout_destroy": (None, [_vp]),
    "h3dmg_layout_error": (ctypes.c_char_p, []),
    "h3dmg_layout_ndomains": (_i, [_vp]),
    "h3dmg_layout_nranks": (_i, [_vp]),
    "h3dmg_layout_domain": (_i, [_vp, _i, _ip, _ip, _dp, _dp]),
    "h3dmg_layout_neighbours": (_i, [_vp, _i, _ip, _ip, _ip]),
    "h3dmg_layout_blocks": (_i, [_vp, _i, _dp, _ip, _i, _P(H3dBlock), _i, _ip, _ip]),
    "h3dmg_get_tail": (_i, [_vp]),
    "h3dmg_velocity_begin": (_i, [_vp, _P(_vp), _P(_vp), _P(_vp), _P(_vp), _P(_vp), _P(_vp), _d,
                                  _dp]),
    "h3dmg_velocity_step": (_i, [_vp, _dp]),
    "h3dmg_solve_velocity": (_i, [_vp, _P(_vp), _P(_vp), _P(_vp), _P(_vp), _P(_vp), _P(_vp), _d,
                                  _dp, _i, _ip, _dp]),
    "h3dmg_get_su_device": (_i, [_vp, _i, _vp]),
    "h3dmg_set_velocity": (_i, [_vp, _i, _dp, _dp, _dp, _dp, _dp, _dp, _i, _i, _i]),
    "h3dmg_get_velocity": (_i, [_vp, _i, _dp, _dp, _dp, _i, _i, _i]),
    "h3dmg_get_su": (_i, [_vp, _i, _dp, _i, _i, _i]),
    "h3dmg_get_velocity_calvel": (_i, [_vp, _i, _dp, _dp, _dp, _i, _i, _i]),
    "h3dmg_set_dt": (_i, [_vp, _d]),
    "h3dmg_set_differencing": (_i, [_vp, _i]),
    "h3dmg_set_density": (_i, [_vp, _i, _dp, _i, _i, _i]),
    "h3dmg_set_density_device": (_i, [_vp, _i, _vp]),
    "h3dmg_tecbin_write": (_i, [ctypes.c_char_p, _P(H3dTecbinHeader), _dp, _dp, _dp, _P(_dp)]),
    "h3dmg_tecbin_read_header": (_i, [ctypes.c_char_p, _P(H3dTecbinHeader)]),
    "h3dmg_tecbin_read": (_i, [ctypes.c_char_p, _i, _dp, _dp, _dp, _P(_dp)]),
    "h3dmg_tecbin_error": (ctypes.c_char_p, []),
    "h3dmg_dbg_put": (_i, [_vp, _i, _i, _i, _vp]),
    "h3dmg_dbg_get": (_i, [_vp, _i, _i, _i, _vp]),
    "h3dmg_dbg_relax": (_i, [_vp, _i, _i]),
    "h3dmg_dbg_restrict": (_i, [_vp, _i]),
    "h3dmg_dbg_prolong": (_i, [_vp, _i]),
}

_lib = None


class H3dError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"h3dmg status {status}: {msg}")
        self.status = status


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: build it with `python -m hydro3d_amd.build`")
        lib = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib


def check(status, allow=()):
    if status != OK and status not in allow:
        raise H3dError(status, load().h3dmg_last_error().decode())
    return status
