"""Hydro3D's multi-block layout files (infodom.cin, mdmap.cin) through the
library's reader (h3dmg_layout_*, hydro3d_amd/csrc/h3dmg_layout.cpp): the
blocks of one rank as MultigridSolver specs with their shared faces.

Reference: read_mdmap / read_infodom and the neighbour and periodic-partner
search, src_v70/alloc_dom.for:284-1215; cells per block from the global
spacing, src_v70/localparameters.for:80-104.
"""
from __future__ import annotations

import ctypes

from . import _lib as L

# control.cin boundary codes: 5 is periodic; every other code is a wall for
# the pressure (src_v70/mgsolver.for:194-205)
PERIODIC_CODE = 5


class Layout:
    def __init__(self, infodom_path, mdmap_path):
        self.lib = L.load()
        h = ctypes.c_void_p()
        st = self.lib.h3dmg_layout_read(str(infodom_path).encode(), str(mdmap_path).encode(),
                                        ctypes.byref(h))
        if st != L.OK:
            raise L.H3dError(st, self.lib.h3dmg_layout_error().decode())
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            self.lib.h3dmg_layout_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    @property
    def ndomains(self) -> int:
        return int(self.lib.h3dmg_layout_ndomains(self.h))

    @property
    def nranks(self) -> int:
        return int(self.lib.h3dmg_layout_nranks(self.h))

    def _check(self, st):
        if st != L.OK:
            raise L.H3dError(st, self.lib.h3dmg_layout_error().decode())

    def domain(self, d):
        """(rank, rdiv, lo[3], hi[3]) of domain d."""
        rank, rdiv = ctypes.c_int(), ctypes.c_int()
        lo, hi = (ctypes.c_double * 3)(), (ctypes.c_double * 3)()
        self._check(self.lib.h3dmg_layout_domain(self.h, d, ctypes.byref(rank), ctypes.byref(rdiv),
                                                 lo, hi))
        return rank.value, rdiv.value, tuple(lo), tuple(hi)

    def neighbours(self, d, bc_codes):
        """(iprev, inext, jprev, jnext, kprev, knext) and the periodic
        partners (per_ip, per_in, per_jp, per_jn, per_kp, per_kn) of domain
        d, -1 for none, under the six control.cin boundary codes."""
        codes = (ctypes.c_int * 6)(*[int(x) for x in bc_codes])
        pn, pp = (ctypes.c_int * 6)(), (ctypes.c_int * 6)()
        self._check(self.lib.h3dmg_layout_neighbours(self.h, d, codes, pn, pp))
        return tuple(pn), tuple(pp)

    def blocks(self, rank, gdx, bc_codes, nlevels):
        """The blocks of `rank` as MultigridSolver specs (bc 0 wall, 1
        periodic onto itself, 2 shared with block nbr[f]) and their domain
        numbers."""
        codes = (ctypes.c_int * 6)(*[int(x) for x in bc_codes])
        g = (ctypes.c_double * 3)(*[float(x) for x in gdx])
        nb = ctypes.c_int()
        self._check(self.lib.h3dmg_layout_blocks(self.h, rank, g, codes, nlevels, None, 0,
                                                 ctypes.byref(nb), None))
        arr = (L.H3dBlock * max(1, nb.value))()
        ids = (ctypes.c_int * max(1, nb.value))()
        self._check(self.lib.h3dmg_layout_blocks(self.h, rank, g, codes, nlevels, arr, nb.value,
                                                 ctypes.byref(nb), ids))
        specs = []
        for q in range(nb.value):
            B = arr[q]
            specs.append(dict(n=(B.ni, B.nj, B.nk), h=(B.dx, B.dy, B.dz), nlev=B.nlevels,
                              bc=tuple(B.bc), nbr=tuple(B.nbr)))
        return specs, [ids[q] for q in range(nb.value)]
