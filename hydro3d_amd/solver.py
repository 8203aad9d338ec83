"""Python mirror of the solver interface (the C-ABI of include/h3dmg.h).

The reference drives the solve through argument-less Fortran subroutines on
module state (`newsolv_mg`, src_v70/newsolv_mg.for:2-93; `mgkcyc`,
src_v70/mgsolver.for:226-372).  `MultigridSolver` holds the same state
explicitly: one handle with every block's multigrid hierarchy resident in HBM.

Array convention: numpy arrays are C-ordered (nk, nj, ni) -- the memory of the
reference's Fortran (i, j, k) arrays.  Device tensors (torch, on the GPU) are
passed by pointer and must be dense float64 of the same shape.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib as L

_PREC = {"fp64": L.FP64, "mixed": L.MIXED}
_COARSE = {"smooth": L.COARSE_SMOOTH, "exact": L.COARSE_EXACT}
_CRIT = {"rel_l2": L.CRIT_REL_L2, "ref_rmax": L.CRIT_REF_RMAX}
# mg_itrsch (src_v70/control.cin): lexicographic GS, line TDMA, red-black GS
_SMOOTH = {"gs": L.SMOOTH_GS, "tdma": L.SMOOTH_TDMA, "rbgs": L.SMOOTH_RBGS}


def _dptr(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _is_device_tensor(x) -> bool:
    return hasattr(x, "data_ptr") and getattr(x, "is_cuda", False)


def _vp(t):
    """Device pointer of tensor t for the library, after torch's current
    stream on t's device has drained.  The library runs on its own
    non-blocking stream, which does not wait for torch's: without this, a
    tensor torch has just computed (or memory torch is still reading) could
    be touched before torch's kernels on it have finished."""
    import torch
    torch.cuda.current_stream(t.device).synchronize()
    return ctypes.c_void_p(t.data_ptr())


def block_array(blocks):
    """h3dmg_block array of block dicts; returns (array, normalised dicts)."""
    nb = len(blocks)
    arr = (L.H3dBlock * nb)()
    norm = []
    for q, b in enumerate(blocks):
        n = tuple(int(x) for x in b["n"])
        arr[q].ni, arr[q].nj, arr[q].nk = n
        arr[q].dx, arr[q].dy, arr[q].dz = (float(x) for x in b["h"])
        arr[q].nlevels = int(b["nlev"])
        arr[q].bc[:] = [int(x) for x in b["bc"]]
        nbr = [int(x) for x in b.get("nbr", (-1,) * 6)]
        arr[q].nbr[:] = nbr
        norm.append(dict(n=n, h=tuple(b["h"]), bc=tuple(b["bc"]), nlev=int(b["nlev"]),
                         nbr=tuple(nbr)))
    return arr, norm


def make_params(nu1=2, nu2=2, maxcy=60, tol=1e-8, criterion="rel_l2", precision="fp64",
                coarse="smooth", min_cycles=1, use_graph=True, dt=1e-3, smoother="rbgs"):
    lib = L.load()
    prm = L.H3dParams()
    lib.h3dmg_default_params(ctypes.byref(prm))
    prm.nu1, prm.nu2, prm.maxcy = nu1, nu2, maxcy
    prm.tol, prm.dt = tol, dt
    prm.criterion = _CRIT[criterion]
    prm.precision = _PREC[precision]
    prm.coarse = _COARSE[coarse]
    prm.smoother = _SMOOTH[smoother]
    prm.min_cycles = min_cycles
    prm.use_graph = 1 if use_graph else 0
    return prm


class DevicePool:
    """h3dmg_pool: one device allocation, cut into the regions of blocks that
    outlive the handles built over them (continuous batching)."""

    def __init__(self, nbytes, max_blocks=4096):
        self.lib = L.load()
        p = ctypes.c_void_p()
        L.check(self.lib.h3dmg_pool_create(int(nbytes), int(max_blocks), ctypes.byref(p)))
        self.p = p
        self.nbytes = int(nbytes)

    def alloc(self, nbytes):
        """Offset of a new piece of `nbytes`; MemoryError when none is free."""
        off = ctypes.c_ulonglong(0)
        st = self.lib.h3dmg_pool_alloc(self.p, int(nbytes), ctypes.byref(off))
        if st == L.ERR_NOMEM:
            raise MemoryError(self.lib.h3dmg_last_error().decode())
        L.check(st)
        return int(off.value)

    def free(self, off):
        L.check(self.lib.h3dmg_pool_free(self.p, int(off)))

    @property
    def free_bytes(self):
        return int(self.lib.h3dmg_pool_free_bytes(self.p))

    @property
    def largest_free(self):
        return int(self.lib.h3dmg_pool_largest_free(self.p))

    def close(self):
        if getattr(self, "p", None):
            self.lib.h3dmg_pool_destroy(self.p)
            self.p = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def block_bytes(spec, **params):
    """Bytes of one block's region in a pool (h3dmg_block_bytes)."""
    lib = L.load()
    arr, _ = block_array([spec])
    prm = make_params(**params)
    n = int(lib.h3dmg_block_bytes(arr, ctypes.byref(prm)))
    if n == 0:
        raise L.H3dError(L.ERR_ARG, lib.h3dmg_last_error().decode())
    return n


class MultigridSolver:
    """Blocks, each given as dict(n=(ni,nj,nk), h=(dx,dy,dz),
    bc=(w,e,s,n,b,t) with 0 wall / 1 periodic / 2 shared with block nbr[f],
    nlev=levels, and for shared faces nbr=(six block indices)).

    With `pool` (a DevicePool) the handle is built in the pool over block
    regions at `offsets` (h3dmg_create_pooled): `fresh[b]` false keeps block
    b's state from an earlier handle over the same region; `stream` is the
    HIP stream pointer to launch on (None: a stream of the handle's own)."""

    def __init__(self, blocks, nu1=2, nu2=2, maxcy=60, tol=1e-8, criterion="rel_l2",
                 precision="fp64", coarse="smooth", min_cycles=1, use_graph=True, dt=1e-3,
                 smoother="rbgs", pool=None, offsets=None, fresh=None, stream=None):
        self.lib = L.load()
        nb = len(blocks)
        arr, self.blocks = block_array(blocks)
        prm = make_params(nu1=nu1, nu2=nu2, maxcy=maxcy, tol=tol, criterion=criterion,
                          precision=precision, coarse=coarse, min_cycles=min_cycles,
                          use_graph=use_graph, dt=dt, smoother=smoother)
        self.params = prm
        h = ctypes.c_void_p()
        if pool is None:
            L.check(self.lib.h3dmg_create(arr, nb, ctypes.byref(prm), ctypes.byref(h)))
        else:
            offs = (ctypes.c_ulonglong * nb)(*[int(o) for o in offsets])
            fr = None if fresh is None else (ctypes.c_int * nb)(*[1 if x else 0 for x in fresh])
            st = self.lib.h3dmg_create_pooled(pool.p, arr, nb, ctypes.byref(prm), offs, fr,
                                              ctypes.c_void_p(stream or 0), ctypes.byref(h))
            if st == L.ERR_NOMEM:
                raise MemoryError(self.lib.h3dmg_last_error().decode())
            L.check(st)
        self.h = h
        self.maxcy = maxcy

    # ---- lifetime
    def close(self):
        if getattr(self, "h", None):
            self.lib.h3dmg_destroy(self.h)
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
    def device_bytes(self) -> int:
        return int(self.lib.h3dmg_device_bytes(self.h))

    def _shape(self, b, pad=0):
        ni, nj, nk = self.blocks[b]["n"]
        return (nk + 2 * pad, nj + 2 * pad, ni + 2 * pad)

    # ---- data in / out
    def set_rhs(self, b, f):
        if _is_device_tensor(f):
            assert tuple(f.shape) == self._shape(b) and f.is_contiguous()
            L.check(self.lib.h3dmg_set_rhs_device(self.h, b, _vp(f)))
            return
        f = np.ascontiguousarray(f, dtype=np.float64)
        assert f.shape == self._shape(b), (f.shape, self._shape(b))
        ni, nj, _ = self.blocks[b]["n"]
        L.check(self.lib.h3dmg_set_rhs(self.h, b, _dptr(f), ni, nj, 0))

    def set_p(self, b, p):
        if _is_device_tensor(p):
            assert tuple(p.shape) == self._shape(b) and p.is_contiguous()
            L.check(self.lib.h3dmg_set_p_device(self.h, b, _vp(p)))
            return
        p = np.ascontiguousarray(p, dtype=np.float64)
        ni, nj, _ = self.blocks[b]["n"]
        if p.shape == self._shape(b, 1):
            L.check(self.lib.h3dmg_set_p(self.h, b, _dptr(p), ni + 2, nj + 2, 1))
        else:
            assert p.shape == self._shape(b)
            L.check(self.lib.h3dmg_set_p(self.h, b, _dptr(p), ni, nj, 0))

    def get_p(self, b, out=None):
        """Pressure with its one-cell ghost layer, shape (nk+2, nj+2, ni+2);
        with a device tensor `out` (nk, nj, ni): the interior, on the GPU."""
        if out is not None and _is_device_tensor(out):
            L.check(self.lib.h3dmg_get_p_device(self.h, b, _vp(out)))
            return out
        ni, nj, _ = self.blocks[b]["n"]
        p = np.zeros(self._shape(b, 1))
        L.check(self.lib.h3dmg_get_p(self.h, b, _dptr(p), ni + 2, nj + 2, 1))
        return p

    def result_p(self, b):
        """The pressure of block b as a new device tensor (nk, nj, ni)."""
        import torch
        out = torch.empty(self._shape(b), dtype=torch.float64, device="cuda")
        return self.get_p(b, out=out)

    def get_pp(self, b):
        ni, nj, _ = self.blocks[b]["n"]
        pp = np.zeros(self._shape(b, 1))
        L.check(self.lib.h3dmg_get_pp(self.h, b, _dptr(pp), ni + 2, nj + 2, 1))
        return pp

    # ---- solver
    def vcycle(self):
        L.check(self.lib.h3dmg_vcycle(self.h))

    def solve(self, raise_on_fail=False):
        """Returns (status, ncycles, hist[0..ncycles]) -- hist is the largest
        relative residual over the blocks (REL_L2) or the global rmax."""
        hist = np.zeros(self.maxcy + 1)
        nc = ctypes.c_int(0)
        fr = ctypes.c_double(0.0)
        st = self.lib.h3dmg_solve(self.h, _dptr(hist), hist.size, ctypes.byref(nc),
                                  ctypes.byref(fr))
        if raise_on_fail:
            L.check(st)
        else:
            L.check(st, allow=(L.ERR_NOT_CONVERGED,))
        return st, nc.value, hist[: nc.value + 1]

    def block_stats(self):
        nb = len(self.blocks)
        cyc = (ctypes.c_int * nb)()
        rel = (ctypes.c_double * nb)()
        L.check(self.lib.h3dmg_block_stats(self.h, cyc, rel))
        return list(cyc), list(rel)

    def project(self, b, ustar, vstar, wstar, alfapr=1.0, want_su=True):
        """calvel + calmas of block b with its current pressure, on the GPU.
        ustar (nk, nj, ni+1), vstar (nk, nj+1, ni), wstar (nk+1, nj, ni):
        device float64 tensors.  Returns (u, v, w, su or None, rmax)."""
        import torch
        for t in (ustar, vstar, wstar):
            assert _is_device_tensor(t) and t.dtype == torch.float64 and t.is_contiguous()
        u, v, w = torch.empty_like(ustar), torch.empty_like(vstar), torch.empty_like(wstar)
        su = torch.empty(self._shape(b), dtype=torch.float64, device=ustar.device) if want_su else None
        rmax = ctypes.c_double(0.0)
        ptr = lambda t: _vp(t) if t is not None else ctypes.c_void_p(0)  # noqa: E731
        L.check(self.lib.h3dmg_project_device(self.h, b, ptr(ustar), ptr(vstar), ptr(wstar),
                                              ptr(u), ptr(v), ptr(w), ptr(su), float(alfapr),
                                              ctypes.byref(rmax)))
        return u, v, w, su, rmax.value

    def _vel_lists(self, ustar, vstar, wstar):
        """Per-block device pointer lists of the velocity-form loop; the
        outputs u, v, w are allocated like the inputs."""
        import torch
        nb = len(self.blocks)
        ins = [list(x) for x in (ustar, vstar, wstar)]
        assert all(len(x) == nb for x in ins)
        outs = [[None if t is None else torch.empty_like(t) for t in x] for x in ins]
        for x in ins:
            for t in x:
                assert t is None or (_is_device_tensor(t) and t.dtype == torch.float64
                                     and t.is_contiguous())
        if any(t is not None for x in ins for t in x):
            torch.cuda.current_stream().synchronize()  # see _vp
        arr = lambda ts: (ctypes.c_void_p * nb)(*[t.data_ptr() if t is not None else 0  # noqa: E731
                                                  for t in ts])
        self._vel_keep = (ins, outs)  # the device arrays must outlive the loop
        return [arr(x) for x in ins], [arr(x) for x in outs], outs

    def solve_velocity(self, ustar=None, vstar=None, wstar=None, alfapr=1.0,
                       raise_on_fail=False):
        """newsolv_mg's velocity-form loop on the device for every block:
        ustar/vstar/wstar are per-block lists of device float64 tensors
        ((nk, nj, ni+1), (nk, nj+1, ni), (nk+1, nj, ni)), or None for the
        arrays set_velocity uploaded; eps is `tol`, dt the solver's dt.
        Returns (status, ncycles, rmax history, (u, v, w) lists or None)."""
        if ustar is None:
            us = vs = ws = u = v = w = None
            outs = None
        else:
            (us, vs, ws), (u, v, w), outs = self._vel_lists(ustar, vstar, wstar)
        hist = np.zeros(self.maxcy + 1)
        nc = ctypes.c_int(0)
        rm = ctypes.c_double(0.0)
        st = self.lib.h3dmg_solve_velocity(self.h, us, vs, ws, u, v, w, float(alfapr), _dptr(hist),
                                           hist.size, ctypes.byref(nc), ctypes.byref(rm))
        if raise_on_fail:
            L.check(st)
        else:
            L.check(st, allow=(L.ERR_NOT_CONVERGED,))
        return st, nc.value, hist[: nc.value + 1], outs

    def set_velocity(self, b, ustar, vstar, wstar, u=None, v=None, w=None, pl=2):
        """Host face arrays in the reference's layout, numpy (nk+2pl, nj+2pl,
        ni+2pl) with u[.., i] the face between cells i and i+1 (0-based
        padded index); u, v, w (optional) supply the faces calvel keeps."""
        ni, nj, _ = self.blocks[b]["n"]
        arrs = [np.ascontiguousarray(a, dtype=np.float64) if a is not None else None
                for a in (ustar, vstar, wstar, u, v, w)]
        star = arrs[:3]
        cur = [a if a is not None else star[q] for q, a in enumerate(arrs[3:])]
        L.check(self.lib.h3dmg_set_velocity(self.h, b, *[_dptr(a) for a in star + cur],
                                            ni + 2 * pl, nj + 2 * pl, pl))

    def get_velocity(self, b, out=None, pl=2):
        """u, v, w of the last calvel into host arrays in the reference's
        layout (zeros where nothing is written unless `out` is given)."""
        ni, nj, nk = self.blocks[b]["n"]
        shape = (nk + 2 * pl, nj + 2 * pl, ni + 2 * pl)
        out = out if out is not None else [np.zeros(shape) for _ in range(3)]
        L.check(self.lib.h3dmg_get_velocity(self.h, b, *[_dptr(a) for a in out],
                                            ni + 2 * pl, nj + 2 * pl, pl))
        return out

    def set_differencing(self, differencing):
        """calmas' branch in the velocity loop: the reference's differencing,
        1 (2nd order) or 2 (4th order, press.for:17-37)."""
        L.check(self.lib.h3dmg_set_differencing(self.h, int(differencing)))

    def get_su_host(self, b, pl=2):
        ni, nj, nk = self.blocks[b]["n"]
        out = np.zeros((nk + 2 * pl, nj + 2 * pl, ni + 2 * pl))
        L.check(self.lib.h3dmg_get_su(self.h, b, _dptr(out), ni + 2 * pl, nj + 2 * pl, pl))
        return out

    def set_density(self, b, dens):
        """Density of block b (interior (nk, nj, ni), numpy or a device
        tensor): the reference's variable-density (L_LSM) operator."""
        if _is_device_tensor(dens):
            assert tuple(dens.shape) == self._shape(b) and dens.is_contiguous()
            L.check(self.lib.h3dmg_set_density_device(self.h, b, _vp(dens)))
            return
        d = np.ascontiguousarray(dens, dtype=np.float64)
        assert d.shape == self._shape(b)
        ni, nj, _ = self.blocks[b]["n"]
        L.check(self.lib.h3dmg_set_density(self.h, b, _dptr(d), ni, nj, 0))

    def set_dt(self, dt):
        L.check(self.lib.h3dmg_set_dt(self.h, float(dt)))

    def velocity_begin(self, ustar, vstar, wstar, alfapr=1.0):
        """First calvel + calmas of the velocity loop; returns (rmax, (u, v, w))."""
        (us, vs, ws), (u, v, w), outs = self._vel_lists(ustar, vstar, wstar)
        rm = ctypes.c_double(0.0)
        L.check(self.lib.h3dmg_velocity_begin(self.h, us, vs, ws, u, v, w, float(alfapr),
                                              ctypes.byref(rm)))
        return rm.value, outs

    def velocity_step(self):
        """One iteration of the velocity loop; returns rmax."""
        rm = ctypes.c_double(0.0)
        L.check(self.lib.h3dmg_velocity_step(self.h, ctypes.byref(rm)))
        return rm.value

    def get_su(self, b):
        """su as the block's rhs arrays hold it, a device tensor (nk, nj, ni)."""
        import torch
        out = torch.empty(self._shape(b), dtype=torch.float64, device="cuda")
        L.check(self.lib.h3dmg_get_su_device(self.h, b, _vp(out)))
        return out

    def set_block_tol(self, b, tol):
        """REL_L2 tolerance of block b (0: the solver's)."""
        L.check(self.lib.h3dmg_set_block_tol(self.h, b, float(tol)))

    def set_block_enabled(self, b, enabled):
        """A disabled block is skipped by solve/vcycle; its state stays."""
        L.check(self.lib.h3dmg_set_block_enabled(self.h, b, 1 if enabled else 0))

    # ---- stepped solves (continuous batching)
    def block_state(self, b):
        st = L.H3dBlockState()
        L.check(self.lib.h3dmg_get_block_state(self.h, b, ctypes.byref(st)))
        return st

    def set_block_state(self, b, st):
        L.check(self.lib.h3dmg_set_block_state(self.h, b, ctypes.byref(st)))

    def solve_begin(self, blocks):
        """Start the stepped solve of the listed blocks (r = f - A p)."""
        idx = (ctypes.c_int * len(blocks))(*[int(b) for b in blocks])
        L.check(self.lib.h3dmg_solve_begin(self.h, idx, len(blocks)))

    def solve_step(self):
        """One outer iteration of every running block; returns how many
        still run.  Each block's outcome is in block_state(b).status."""
        n = ctypes.c_int(0)
        L.check(self.lib.h3dmg_solve_step(self.h, ctypes.byref(n)))
        return n.value

    def remove_mean(self):
        L.check(self.lib.h3dmg_remove_mean(self.h))

    def set_stream(self, stream_ptr: int | None):
        L.check(self.lib.h3dmg_set_stream(self.h, ctypes.c_void_p(stream_ptr or 0)))

    # ---- tuning
    def set_march(self, level, planes):
        L.check(self.lib.h3dmg_set_march(self.h, level, planes))

    def get_march(self, level):
        return int(self.lib.h3dmg_get_march(self.h, level))

    def set_lanes(self, level, lanes):
        L.check(self.lib.h3dmg_set_lanes(self.h, level, lanes))

    def get_lanes(self, level):
        return int(self.lib.h3dmg_get_lanes(self.h, level))

    def set_fused(self, max_cells):
        """Levels of at most max_cells cells per block: one fused launch per
        red-black sweep (0 = never); results are identical either way."""
        L.check(self.lib.h3dmg_set_fused(self.h, int(max_cells)))

    def get_fused(self):
        return int(self.lib.h3dmg_get_fused(self.h))

    def set_tail(self, level):
        """Levels >= level run in the single-workgroup V-cycle tail (0 = auto,
        above the coarsest = off); results are identical either way."""
        L.check(self.lib.h3dmg_set_tail(self.h, level))

    def get_tail(self):
        return int(self.lib.h3dmg_get_tail(self.h))

    @property
    def nlevels_max(self):
        return max(b["nlev"] for b in self.blocks)

    # ---- operator seams (device tensors in the oracle's level layout)
    def dbg_put(self, b, level, which, t):
        L.check(self.lib.h3dmg_dbg_put(self.h, b, level, which, _vp(t)))

    def dbg_get(self, b, level, which, t):
        L.check(self.lib.h3dmg_dbg_get(self.h, b, level, which, _vp(t)))
        return t

    def dbg_relax(self, level, nsweep):
        L.check(self.lib.h3dmg_dbg_relax(self.h, level, nsweep))

    def dbg_restrict(self, level):
        L.check(self.lib.h3dmg_dbg_restrict(self.h, level))

    def dbg_prolong(self, level):
        L.check(self.lib.h3dmg_dbg_prolong(self.h, level))
