"""Blocks with neighbours on the device against the reference build.
Needs an MI355X.

The fixtures (tests/golden/mb_*.npz, written by make_golden_multiblock.py)
hold the reference's own multi-block run: its layout reader and neighbour
search (alloc_dom.for), exchangepp/exchbc_mgpp between the blocks of one
rank, and the multigrid and velocity loop over them.  The device holds no
ghosts: every ghost a kernel reads is the neighbour's cell that exchangepp
would have copied.  Every comparison is bitwise, the ghost layers included
(after relax: the state exchangepp leaves; after mgcorr: the corrected
neighbour values, which mgbound alone leaves in the reference's ghosts).
"""
import glob
import os

import numpy as np
import pytest

from h3d_util import GOLDEN, load_golden

pytestmark = pytest.mark.gpu

MB = sorted(p for p in glob.glob(os.path.join(GOLDEN, "mb_*.npz"))
            if "benchmark" not in os.path.basename(p))


def _dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).cuda()


def specs_from_table(z):
    """Block specs from the reference's neighbour table (iprev..knext and
    the periodic partners per_ip..per_kn): a shared face is NEIGHBOUR, a
    periodic axis spanned by one block PERIODIC, the rest walls."""
    tab, bc = z["nbr_table"], tuple(int(x) for x in z["bc"])
    specs = []
    for b, row in enumerate(tab):
        prevnext, partner = row[4:10], row[30:36]
        bcs, nbr = [], []
        for f in range(6):
            if prevnext[f] >= 0:
                bcs.append(2)
                nbr.append(int(prevnext[f]))
            elif bc[2 * (f // 2)] == 1:
                if prevnext[f ^ 1] < 0:
                    bcs.append(1)
                    nbr.append(-1)
                else:
                    bcs.append(2)
                    nbr.append(int(partner[f]))
            else:
                bcs.append(0)
                nbr.append(-1)
        specs.append(dict(n=tuple(int(x) for x in row[1:4]), h=tuple(float(x) for x in z["h"]),
                          bc=tuple(bcs), nbr=tuple(nbr), nlev=int(z["nlev"])))
    return specs


def _solver(specs, **kw):
    from hydro3d_amd import MultigridSolver
    return MultigridSolver(specs, **kw)


def _shape(s, g, pad):
    ni, nj, nk = (x >> (g - 1) for x in s["n"])
    return (nk + 2 * pad, nj + 2 * pad, ni + 2 * pad)


def _put(S, b, g, which, interior):
    from hydro3d_amd import _lib as L
    a = np.pad(interior, 1) if which == L.ARR_PHI else interior
    S.dbg_put(b, g, which, _dev(a))


def _get(S, b, g, which, shape):
    import torch
    return S.dbg_get(b, g, which, torch.empty(shape, dtype=torch.float64, device="cuda")).cpu().numpy()


@pytest.mark.parametrize("path", MB, ids=[os.path.basename(p) for p in MB])
def test_gpu_multiblock_operators_bitwise_vs_reference(gpu_lib, path):
    from hydro3d_amd import _lib as L
    z = load_golden(os.path.basename(path))
    specs = specs_from_table(z)
    nb, Lv = len(specs), int(z["nlev"])
    print(f"case {os.path.basename(path)}: {nb} blocks, {Lv} levels", flush=True)
    with _solver(specs, use_graph=False) as S:
        for g in range(1, Lv + 1):
            for b in range(nb):
                _put(S, b, g, L.ARR_PHI, z[f"relax{g}_in_phi_{b}"])
                _put(S, b, g, L.ARR_RHS, z[f"relax{g}_in_rhs_{b}"])
            S.dbg_relax(g, 2)
            for b in range(nb):
                out = _get(S, b, g, L.ARR_PHI, _shape(specs[b], g, 1))
                ref = z[f"relax{g}_out_{b}"]
                assert np.array_equal(out[1:-1, 1:-1, 1:-1], ref[1:-1, 1:-1, 1:-1]), \
                    f"relax level {g} block {b} interior"
                assert np.array_equal(out, ref), f"relax level {g} block {b} ghosts"
            if g > 1:
                for b in range(nb):
                    _put(S, b, g - 1, L.ARR_PHI, z[f"restr{g}_in_phi_{b}"])
                    _put(S, b, g - 1, L.ARR_RHS, z[f"restr{g}_in_rhs_{b}"])
                S.dbg_restrict(g)
                for b in range(nb):
                    out = _get(S, b, g, L.ARR_RHS, _shape(specs[b], g, 0))
                    assert np.array_equal(out, z[f"restr{g}_out_rhs_{b}"]), \
                        f"restriction into level {g} block {b}"
            if g < Lv:
                for b in range(nb):
                    _put(S, b, g, L.ARR_PHI, z[f"corr{g}_in_fine_{b}"])
                    _put(S, b, g + 1, L.ARR_PHI, z[f"corr{g}_in_coarse_{b}"])
                S.dbg_prolong(g)
                for b in range(nb):
                    out = _get(S, b, g, L.ARR_PHI, _shape(specs[b], g, 1))
                    ref = z[f"corr{g}_out_{b}"]
                    assert np.array_equal(out[1:-1, 1:-1, 1:-1], ref[1:-1, 1:-1, 1:-1]), \
                        f"prolongation level {g} block {b} interior"
                    assert np.array_equal(out, ref), f"prolongation level {g} block {b} ghosts"


@pytest.mark.parametrize("graph", [True, False], ids=["graph", "eager"])
@pytest.mark.parametrize("path", MB, ids=[os.path.basename(p) for p in MB])
def test_gpu_multiblock_vcycle_bitwise_vs_reference(gpu_lib, path, graph):
    z = load_golden(os.path.basename(path))
    specs = specs_from_table(z)
    print(f"case {os.path.basename(path)} vcycle graph={graph}", flush=True)
    with _solver(specs, use_graph=graph) as S:
        for b in range(len(specs)):
            S.set_rhs(b, z[f"vcycle_su_{b}"])
        S.vcycle()
        for b in range(len(specs)):
            assert np.array_equal(S.get_pp(b), z[f"vcycle_pp_{b}"]), f"block {b}"


def _faces(z, specs):
    from hydro3d_amd import configs as C
    lo = z["lo"]
    h = tuple(float(x) for x in z["h"])
    N = [0, 0, 0]
    for b, s in enumerate(specs):
        for a in range(3):
            N[a] = max(N[a], int(round(lo[b][a] / h[a])) + s["n"][a])
    u, v, w = C.face_velocity(tuple(N), tuple(int(x) for x in z["bc"]), int(z["newsolv_seed"]),
                              0.0, h)
    out = []
    for b, s in enumerate(specs):
        x0, y0, z0 = (int(round(lo[b][a] / h[a])) for a in range(3))
        ni, nj, nk = s["n"]
        out.append((u[z0:z0 + nk, y0:y0 + nj, x0:x0 + ni + 1],
                    v[z0:z0 + nk, y0:y0 + nj + 1, x0:x0 + ni],
                    w[z0:z0 + nk + 1, y0:y0 + nj, x0:x0 + ni]))
    return out


@pytest.mark.parametrize("graph", [True, False], ids=["graph", "eager"])
@pytest.mark.parametrize("path", MB, ids=[os.path.basename(p) for p in MB])
def test_gpu_multiblock_velocity_loop_bitwise_vs_reference(gpu_lib, path, graph):
    """newsolv_mg's loop over blocks with neighbours on the device: calvel
    (the shared faces computed by the block before them and copied, as
    exchange(1..3) does), calmas, mgkcyc with the halo exchange on every
    level, p = p + pp: rmax of every iteration, every block's p with its
    ghost layer and su, bitwise the reference build's."""
    z = load_golden(os.path.basename(path))
    if "newsolv_rmax" not in z:
        pytest.skip("no velocity-loop fixture for this layout")
    specs = specs_from_table(z)
    nc = int(z["newsolv_ncycles"])
    faces = _faces(z, specs)
    print(f"case {os.path.basename(path)} velocity loop graph={graph}", flush=True)
    vel = [[_dev(f[a]) for f in faces] for a in range(3)]
    with _solver(specs, maxcy=nc, tol=0.0, dt=1e-3, use_graph=graph) as S:
        for b, s in enumerate(specs):
            S.set_p(b, np.zeros(_shape(s, 1, 0)))
        st, ran, hist, _ = S.solve_velocity(*vel)
        ps = [S.get_p(b) for b in range(len(specs))]
        sus = [S.get_su(b).cpu().numpy() for b in range(len(specs))]
    assert ran == nc
    assert np.array_equal(hist, z["newsolv_rmax"]), (hist, z["newsolv_rmax"])
    for b in range(len(specs)):
        assert np.array_equal(ps[b], z[f"newsolv_p_{b}"]), f"p of block {b}"
        assert np.array_equal(sus[b], z[f"newsolv_su_{b}"]), f"su of block {b}"
    # with eps = 1e-6 the loop stops where the reference's stops
    with _solver(specs, maxcy=nc, tol=1e-6, dt=1e-3, use_graph=graph) as S:
        for b, s in enumerate(specs):
            S.set_p(b, np.zeros(_shape(s, 1, 0)))
        st, ran, hist, _ = S.solve_velocity(*vel, raise_on_fail=True)
    assert ran == int(z["newsolv_stop_at"])
