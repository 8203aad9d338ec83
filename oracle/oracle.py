"""ctypes wrapper of the C restatement (liboracle.so) and of the reference
Fortran driver (_ref/ref_driver).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
CPU-baseline leg of bench.py as the checker; the product path never imports
it.

Array convention: numpy arrays are C-ordered with shape (nk, nj, ni) -- the
same memory as a Fortran (i, j, k) array.  Padded level arrays carry one
ghost layer: shape (nk+2, nj+2, ni+2).
"""
from __future__ import annotations

import ctypes
import os
import subprocess
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_DRIVER = os.path.join(HERE, "_ref", "ref_driver")

BC_WALL = 0
BC_PERIODIC = 1
# reference bc codes (control.cin): 5 = periodic, anything else = Neumann wall
_REF_BC = {BC_WALL: 1, BC_PERIODIC: 5}

SCHEME_GS, SCHEME_TDMA, SCHEME_RB = 1, 2, 3


class _Block(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int * 3), ("h", ctypes.c_double * 3),
                ("bc", ctypes.c_int * 6), ("nlev", ctypes.c_int)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle library missing: {LIB_PATH} (run `make -C oracle`)")
        L = ctypes.CDLL(LIB_PATH)
        P = ctypes.POINTER
        dp = P(ctypes.c_double)
        bp = P(_Block)
        L.or_relax_rb.argtypes = [bp, ctypes.c_int, ctypes.c_int, dp, dp]
        L.or_relax_gs.argtypes = [bp, ctypes.c_int, ctypes.c_int, dp, dp]
        L.or_relax_tdma.argtypes = [bp, ctypes.c_int, ctypes.c_int, dp, dp]
        L.or_bound.argtypes = [bp, ctypes.c_int, dp]
        L.or_exchange.argtypes = [bp, ctypes.c_int, dp]
        L.or_residual.argtypes = [bp, ctypes.c_int, dp, dp, dp]
        L.or_restrict.argtypes = [bp, ctypes.c_int, dp, dp, dp, dp]
        L.or_prolong.argtypes = [bp, ctypes.c_int, dp, dp]
        L.or_vcycle.argtypes = [bp, ctypes.c_int, ctypes.c_int, ctypes.c_int, dp, dp]
        L.or_vcycle.restype = ctypes.c_int
        L.or_solve.argtypes = [bp, ctypes.c_int, ctypes.c_int, ctypes.c_int, dp, dp,
                               ctypes.c_int, ctypes.c_double, dp]
        L.or_solve.restype = ctypes.c_int
        L.or_apply.argtypes = [bp, dp, dp]
        _lib = L
    return _lib


def _dp(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


class Block:
    """One multigrid block: interior size n=(ni,nj,nk), spacing h, six face
    boundary kinds (west, east, south, north, bottom, top), level count."""

    def __init__(self, n, h, bc, nlev):
        self.n = tuple(int(x) for x in n)
        self.h = tuple(float(x) for x in h)
        self.bc = tuple(int(x) for x in bc)
        self.nlev = int(nlev)

    def c(self):
        b = _Block()
        b.n[:] = self.n
        b.h[:] = self.h
        b.bc[:] = self.bc
        b.nlev = self.nlev
        return ctypes.byref(b), b

    def dims(self, g):
        return tuple(x >> (g - 1) for x in self.n)

    def padded_shape(self, g):
        ni, nj, nk = self.dims(g)
        return (nk + 2, nj + 2, ni + 2)

    def interior_shape(self, g):
        ni, nj, nk = self.dims(g)
        return (nk, nj, ni)


def _call(fn, block, *args):
    ref, keep = block.c()
    return fn(ref, *args)


_RELAX = {SCHEME_GS: "or_relax_gs", SCHEME_TDMA: "or_relax_tdma", SCHEME_RB: "or_relax_rb"}


def relax(block, g, nsweep, phi, rhs, scheme=SCHEME_RB):
    phi = np.ascontiguousarray(phi, dtype=np.float64).copy()
    rhs = np.ascontiguousarray(rhs, dtype=np.float64)
    _call(getattr(lib(), _RELAX[scheme]), block, g, nsweep, _dp(phi), _dp(rhs))
    return phi


def exchange(block, g, phi):
    phi = np.ascontiguousarray(phi, dtype=np.float64).copy()
    _call(lib().or_exchange, block, g, _dp(phi))
    return phi


def bound(block, g, phi):
    phi = np.ascontiguousarray(phi, dtype=np.float64).copy()
    _call(lib().or_bound, block, g, _dp(phi))
    return phi


def residual(block, g, phi, rhs):
    phi = np.ascontiguousarray(phi, dtype=np.float64)
    rhs = np.ascontiguousarray(rhs, dtype=np.float64)
    res = np.zeros(block.interior_shape(g))
    _call(lib().or_residual, block, g, _dp(phi), _dp(rhs), _dp(res))
    return res


def restrict(block, g, phi_f, rhs_f):
    phi_f = np.ascontiguousarray(phi_f, dtype=np.float64)
    rhs_f = np.ascontiguousarray(rhs_f, dtype=np.float64)
    phi_c = np.full(block.padded_shape(g), 7.0)
    rhs_c = np.full(block.interior_shape(g), 7.0)
    _call(lib().or_restrict, block, g, _dp(phi_f), _dp(rhs_f), _dp(phi_c), _dp(rhs_c))
    return phi_c, rhs_c


def prolong(block, g, phi_f, phi_c):
    phi_f = np.ascontiguousarray(phi_f, dtype=np.float64).copy()
    phi_c = np.ascontiguousarray(phi_c, dtype=np.float64)
    _call(lib().or_prolong, block, g, _dp(phi_f), _dp(phi_c))
    return phi_f


def vcycle(block, su, nu1=2, nu2=2, scheme=SCHEME_RB, pp=None):
    su = np.ascontiguousarray(su, dtype=np.float64)
    pp = np.zeros(block.padded_shape(1)) if pp is None else np.array(pp, dtype=np.float64)
    rc = _call(lib().or_vcycle, block, scheme, nu1, nu2, _dp(su), _dp(pp))
    if rc != 0:
        raise ValueError(f"or_vcycle failed ({rc})")
    return pp


def solve(block, f, ncycles, tol=0.0, nu1=2, nu2=2, scheme=SCHEME_RB, p0=None):
    """Pressure-form outer loop; returns (p_padded, hist[0..ran])."""
    f = np.ascontiguousarray(f, dtype=np.float64)
    p = np.zeros(block.padded_shape(1)) if p0 is None else np.array(p0, dtype=np.float64)
    hist = np.zeros(ncycles + 1)
    rc = _call(lib().or_solve, block, scheme, nu1, nu2, _dp(f), _dp(p), ncycles, tol, _dp(hist))
    if rc < 0:
        raise ValueError(f"or_solve failed ({rc})")
    return p, hist[: rc + 1]


def apply(block, p):
    p = np.ascontiguousarray(p, dtype=np.float64)
    out = np.zeros(block.interior_shape(1))
    _call(lib().or_apply, block, _dp(p), _dp(out))
    return out


# ---------------------------------------------------------------------------
# Reference Fortran driver (built from the reference sources into _ref/).

def ref_available() -> bool:
    return os.path.exists(REF_DRIVER)


def run_ref(block, mode, inputs, nout, nu1=2, nu2=2, level=1, nsweep=1, ncycles=1,
            scheme=SCHEME_RB, dt=1e-3, eps=1e-6, timeout=600, lsm=False, cds4=False):
    """Run the reference driver; `inputs` is a list of arrays written in order
    (with lsm, the interior density comes first: the reference's L_LSM
    variable-density operator; with cds4, differencing = 2 and the velocity
    faces of every axis run -1..n+1); returns the flat float64 output
    (length nout, or all if nout is None)."""
    if not ref_available():
        raise RuntimeError(f"reference driver not built: {REF_DRIVER}")
    bcs = ",".join(str(_REF_BC[x]) for x in block.bc)
    with tempfile.TemporaryDirectory() as td:
        job = os.path.join(td, "job.nml")
        fin = os.path.join(td, "in.bin")
        fout = os.path.join(td, "out.bin")
        with open(job, "w") as fh:
            fh.write(
                "&job\n"
                f" mode='{mode}', ni={block.n[0]}, nj={block.n[1]}, nk={block.n[2]},\n"
                f" nlev={block.nlev}, nu1={nu1}, nu2={nu2}, level={level}, nsw={nsweep},\n"
                f" ncycles={ncycles}, bcs={bcs}, h={block.h[0]!r},{block.h[1]!r},{block.h[2]!r},\n"
                f" dtin={dt!r}, epsin={eps!r}, scheme={scheme}, lsm={1 if lsm else 0},\n"
                f" cds4={1 if cds4 else 0}\n/\n")
        with open(fin, "wb") as fh:
            for a in inputs:
                fh.write(np.ascontiguousarray(a, dtype="<f8").tobytes())
        subprocess.run([REF_DRIVER, job, fin, fout], check=True, timeout=timeout,
                       stdout=subprocess.DEVNULL)
        out = np.fromfile(fout, dtype="<f8")
    if nout is not None and out.size != nout:
        raise RuntimeError(f"reference driver wrote {out.size} doubles, expected {nout}")
    return out


# ---------------------------------------------------------------------------
# Several blocks on one rank: the reference's own layout files and neighbour
# search (alloc_dom.for) drive the reference multigrid (_ref/ref_multiblock).

REF_MULTIBLOCK = os.path.join(HERE, "_ref", "ref_multiblock")


def multiblock_available() -> bool:
    return os.path.exists(REF_MULTIBLOCK)


def box_layout(xs, ys, zs, gdx):
    """A box of blocks: per-axis lists of block cell counts (a Cartesian
    product, x fastest in the domain numbering, as the reference's type-2
    infodom numbers them, alloc_dom.for:470-481).  Returns a list of domain
    dicts: id, n (cells), lo/hi coordinates (multiples of gdx)."""
    doms = []
    z0 = 0
    for nz in zs:
        y0 = 0
        for ny in ys:
            x0 = 0
            for nx in xs:
                lo = (x0 * gdx[0], y0 * gdx[1], z0 * gdx[2])
                hi = ((x0 + nx) * gdx[0], (y0 + ny) * gdx[1], (z0 + nz) * gdx[2])
                doms.append(dict(id=len(doms), n=(nx, ny, nz), lo=lo, hi=hi))
                x0 += nx
            y0 += ny
        z0 += nz
    return doms


def infodom_text(doms):
    """infodom.cin, type 1 (per-domain coordinates, alloc_dom.for:442-461)."""
    lines = [" 1  (1)sub-domains description; (2)entire domain description (Lx,Ly,Lz)",
             f"  {len(doms)} number of domains",
             "========== Domain description ==============",
             "0.0  Lx", "0.0  Ly", "0.0  Lz", "  1  nx", "  1  ny", "  1  nz",
             "== CPU number, Local mesh refinement level,xini,xend,yini,yend,zini,zend=="]
    for d in doms:
        c = " ".join(repr(float(v)) for a in range(3) for v in (d["lo"][a], d["hi"][a]))
        lines.append(f"  {d['id']}  1  {c}")
    return "\n".join(lines) + "\n"


def mdmap_text(ndoms, order=None):
    """mdmap.cin, type 1, every domain on rank 0 (alloc_dom.for:301-345)."""
    order = list(range(ndoms)) if order is None else list(order)
    return (" 1  (1)sub-domains description\n"
            f"  {ndoms} number of domain\n"
            "  1 number of processor\n"
            " ===============================================\n"
            f"  0  {len(order)}  " + " ".join(str(x) for x in order) + "\n")


def run_multiblock(doms, mode, inputs, nout, gdx, bcs, **kw):
    """Run the multi-block reference driver on the layout `doms` (one rank;
    blocks in domain order); `bcs` are the six global boundary kinds
    (BC_WALL / BC_PERIODIC) as control.cin sets them for every block."""
    return run_multiblock_files(infodom_text(doms), mdmap_text(len(doms)), mode, inputs, nout,
                                gdx, bcs, **kw)


DROPIN_DRIVER = os.path.join(HERE, "_ref", "dropin_driver")


def dropin_available() -> bool:
    return os.path.exists(DROPIN_DRIVER)


def mdmap_ranks_text(ndoms, nranks):
    """mdmap.cin, type 1, the domains dealt out in contiguous runs over
    `nranks` ranks (alloc_dom.for:301-345)."""
    per = [ndoms // nranks + (1 if r < ndoms % nranks else 0) for r in range(nranks)]
    lines = [" 1  (1)sub-domains description", f"  {ndoms} number of domain",
             f"  {nranks} number of processor", " ==============================================="]
    d = 0
    for r, m in enumerate(per):
        lines.append(f"  {r}  {m}  " + " ".join(str(x) for x in range(d, d + m)))
        d += m
    return "\n".join(lines) + "\n"


def block_inputs(doms, u, v, w, gdx):
    """Per domain u*, v*, w* (faces 0..n of each axis) cut from the face
    arrays of the whole domain, and p = 0, in domain order."""
    ins = []
    for d in doms:
        x0, y0, z0 = (int(round(d["lo"][a] / gdx[a])) for a in range(3))
        ni, nj, nk = d["n"]
        ins += [u[z0:z0 + nk, y0:y0 + nj, x0:x0 + ni + 1], v[z0:z0 + nk, y0:y0 + nj + 1, x0:x0 + ni],
                w[z0:z0 + nk + 1, y0:y0 + nj, x0:x0 + ni], np.zeros((nk, nj, ni))]
    return ins


def time_multiblock(doms, inputs, gdx, bcs, nlev, ncycles, nranks, nu1=2, nu2=2, dt=1e-3,
                    timeout=900, mpiexec="/opt/conda/bin/mpiexec"):
    """The reference's newsolv_mg loop on `nranks` MPI ranks (the blocks of
    `doms`, all of one size, dealt out over the ranks), `ncycles`
    iterations: returns (seconds of the loop on rank 0 between barriers,
    last rmax).  `inputs`: per domain u*, v*, w*, p0 in domain order."""
    if not os.path.exists(REF_MULTIBLOCK):
        raise RuntimeError(f"driver not built: {REF_MULTIBLOCK}")
    codes = ",".join(str(_REF_BC[x]) for x in bcs)
    with tempfile.TemporaryDirectory() as td:
        with open(os.path.join(td, "infodom.cin"), "w") as fh:
            fh.write(infodom_text(doms))
        with open(os.path.join(td, "mdmap.cin"), "w") as fh:
            fh.write(mdmap_ranks_text(len(doms), nranks))
        with open(os.path.join(td, "job.nml"), "w") as fh:
            fh.write("&job\n"
                     f" mode='timeloop', nlev={nlev}, nu1={nu1}, nu2={nu2}, ncycles={ncycles},\n"
                     f" bcs={codes}, gdx={gdx[0]!r},{gdx[1]!r},{gdx[2]!r}, dtin={dt!r}\n/\n")
        with open(os.path.join(td, "in.bin"), "wb") as fh:
            for a in inputs:
                fh.write(np.ascontiguousarray(a, dtype="<f8").tobytes())
        subprocess.run([mpiexec, "-n", str(nranks), REF_MULTIBLOCK, "job.nml", "in.bin", "out.bin"],
                       check=True, timeout=timeout, cwd=td, stdout=subprocess.DEVNULL)
        out = np.fromfile(os.path.join(td, "out.bin"), dtype="<f8")
    return float(out[0]), float(out[2])


def run_multiblock_files(infodom, mdmap, mode, inputs, nout, gdx, bcs, nlev=1, nu1=2, nu2=2,
                         level=1, nsweep=1, ncycles=1, scheme=SCHEME_RB, dt=1e-3, eps=1e-6,
                         timeout=600, early=0, driver=None, cds4=False):
    """The same from the texts of infodom.cin and mdmap.cin (one rank).
    `driver` = DROPIN_DRIVER runs the product's drop-in newsolv_mg on the
    same module state instead of the reference's loop (mode 'newsolv')."""
    driver = driver or REF_MULTIBLOCK
    if not os.path.exists(driver):
        raise RuntimeError(f"driver not built: {driver}")
    codes = ",".join(str(_REF_BC[x]) for x in bcs)
    with tempfile.TemporaryDirectory() as td:
        with open(os.path.join(td, "infodom.cin"), "w") as fh:
            fh.write(infodom)
        with open(os.path.join(td, "mdmap.cin"), "w") as fh:
            fh.write(mdmap)
        with open(os.path.join(td, "job.nml"), "w") as fh:
            fh.write("&job\n"
                     f" mode='{mode}', nlev={nlev}, nu1={nu1}, nu2={nu2}, level={level},\n"
                     f" nsw={nsweep}, ncycles={ncycles}, bcs={codes},\n"
                     f" gdx={gdx[0]!r},{gdx[1]!r},{gdx[2]!r}, dtin={dt!r}, epsin={eps!r},\n"
                     f" scheme={scheme}, early={early}, cds4={1 if cds4 else 0}\n/\n")
        with open(os.path.join(td, "in.bin"), "wb") as fh:
            for a in inputs:
                fh.write(np.ascontiguousarray(a, dtype="<f8").tobytes())
        subprocess.run([driver, "job.nml", "in.bin", "out.bin"], check=True,
                       timeout=timeout, cwd=td, stdout=subprocess.DEVNULL)
        out = np.fromfile(os.path.join(td, "out.bin"), dtype="<f8")
    if nout is not None and out.size != nout:
        raise RuntimeError(f"multi-block driver wrote {out.size} doubles, expected {nout}")
    return out
