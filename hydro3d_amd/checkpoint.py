"""Hydro3D's tecbin restart files through the library's reader and writer
(h3dmg_tecbin_*, hydro3d_amd/csrc/h3dmg_tecbin.cpp), and the warm start of a
solve from one.

Reference: the writer `tecbin`, src_v70/post.for:324-375; the restart
reader, src_v70/initial.for:715-747, which takes the pressure (column 4)
back into dom(ib)%p with its ghost layers.

Arrays are numpy (ttc_k, ttc_j, ttc_i) float64 -- the memory of the
reference's Fortran (ttc_i, ttc_j, ttc_k) arrays, ghost layers included.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib as L

FIELDS = L.TECBIN_FIELDS


def _dptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _check(st):
    if st != L.OK:
        raise L.H3dError(st, L.load().h3dmg_tecbin_error().decode())


def write_tecbin(path, pl, ind, x, y, z, fields):
    """Write one block's restart file.  `fields`: dict name -> array of shape
    (ttc_k, ttc_j, ttc_i) for every name in FIELDS (missing names are
    written as zeros); x, y, z the cell coordinates (ttc_i, ttc_j, ttc_k)."""
    lib = L.load()
    x, y, z = (np.ascontiguousarray(a, dtype=np.float64) for a in (x, y, z))
    shape = (z.size, y.size, x.size)
    hd = L.H3dTecbinHeader()
    hd.tt[0], hd.tt[1], hd.tt[2] = x.size, y.size, z.size
    hd.pl = int(pl)
    for a in range(3):
        hd.ind[a] = int(ind[a])
    zero = np.zeros(shape)
    arrs = []
    for name in FIELDS:
        a = np.ascontiguousarray(fields.get(name, zero), dtype=np.float64)
        if a.shape != shape:
            raise ValueError(f"field {name}: shape {a.shape}, expected {shape}")
        arrs.append(a)
    ptrs = (ctypes.POINTER(ctypes.c_double) * len(arrs))(*[_dptr(a) for a in arrs])
    _check(lib.h3dmg_tecbin_write(str(path).encode(), ctypes.byref(hd), _dptr(x), _dptr(y),
                                  _dptr(z), ptrs))


def read_header(path):
    hd = L.H3dTecbinHeader()
    _check(L.load().h3dmg_tecbin_read_header(str(path).encode(), ctypes.byref(hd)))
    return dict(tt=tuple(hd.tt), pl=hd.pl, ind=tuple(hd.ind))


def read_tecbin(path, pl=-1, names=FIELDS):
    """(header, x, y, z, {name: array}) of a restart file; pl >= 0 refuses a
    file written with another number of ghost layers, as initial.for does."""
    hdr = read_header(path)
    ti, tj, tk = hdr["tt"]
    x, y, z = np.empty(ti), np.empty(tj), np.empty(tk)
    out = {n: np.empty((tk, tj, ti)) for n in names}
    ptrs = (ctypes.POINTER(ctypes.c_double) * len(FIELDS))(
        *[_dptr(out[n]) if n in out else ctypes.POINTER(ctypes.c_double)() for n in FIELDS])
    _check(L.load().h3dmg_tecbin_read(str(path).encode(), int(pl), _dptr(x), _dptr(y), _dptr(z),
                                      ptrs))
    return hdr, x, y, z, out


def warm_start(solver, b, path, pl=None):
    """Set block b's pressure from a restart file (the interior of column 4,
    initial.for:731): the warm start of the next solve.  The file's interior
    must match the block's size."""
    hdr = read_header(path)
    pl = hdr["pl"] if pl is None else pl
    _, _, _, _, f = read_tecbin(path, pl=pl, names=("p",))
    p = f["p"]
    n = tuple(solver.blocks[b]["n"])
    inner = tuple(t - 2 * pl for t in hdr["tt"])
    if inner != n:
        raise ValueError(f"restart file holds {inner} interior cells, block {b} has {n}")
    solver.set_p(b, np.ascontiguousarray(p[pl:pl + n[2], pl:pl + n[1], pl:pl + n[0]]))
    return hdr
