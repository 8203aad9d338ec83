"""The lexicographic Gauss-Seidel and line-TDMA smoothers (mg_itrsch = 1 and
2, src_v70/mgsolver.for:407-549) on the device, through the C-ABI.  Needs an
MI355X.

Bar: bitwise.  One sweep per level and one V-cycle against the reference
build's fixtures (constant and variable-density coefficients); whole solves
of unequal blocks in one handle against the oracle, in a captured graph and
eagerly.
"""
import glob
import os

import numpy as np
import pytest

from h3d_util import GOLDEN, load_golden

pytestmark = pytest.mark.gpu

OPS = sorted(glob.glob(os.path.join(GOLDEN, "ops_*.npz")))
LSM_LINE = sorted(glob.glob(os.path.join(GOLDEN, "lsm_line_*.npz")))
SMOOTHERS = [("gs", 1), ("tdma", 2)]


def _blockspec(z):
    return dict(n=tuple(int(x) for x in z["n"]), h=tuple(float(x) for x in z["h"]),
                bc=tuple(int(x) for x in z["bc"]), nlev=int(z["nlev"]))


def _dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).cuda()


def _solver(specs, **kw):
    from hydro3d_amd import MultigridSolver
    return MultigridSolver(specs if isinstance(specs, list) else [specs], **kw)


def _relax_and_vcycle(S, z, scheme, Lv):
    import torch
    from hydro3d_amd import _lib as L
    for g in range(1, Lv + 1):
        phi, rhs = z[f"relax{g}_s{scheme}_in_phi"], z[f"relax{g}_s{scheme}_in_rhs"]
        S.dbg_put(0, g, L.ARR_PHI, _dev(phi))
        S.dbg_put(0, g, L.ARR_RHS, _dev(rhs))
        S.dbg_relax(g, 1)
        out = S.dbg_get(0, g, L.ARR_PHI, torch.empty(phi.shape, dtype=torch.float64,
                                                     device="cuda")).cpu().numpy()
        assert np.array_equal(out, z[f"relax{g}_s{scheme}_out"]), f"relax level {g}"
    S.set_rhs(0, z[f"vcycle_s{scheme}_su"])
    S.vcycle()
    assert np.array_equal(S.get_pp(0), z[f"vcycle_s{scheme}_pp"]), "vcycle"


@pytest.mark.parametrize("smoother,scheme", SMOOTHERS, ids=[s for s, _ in SMOOTHERS])
@pytest.mark.parametrize("path", OPS, ids=[os.path.basename(p) for p in OPS])
def test_gpu_line_smoothers_bitwise_vs_reference(gpu_lib, path, smoother, scheme):
    """One GS / TDMA sweep of every level from the reference's inputs, and a
    V-cycle (nu1 = nu2 = 2) from its su, are bitwise the reference build's
    (ghosts included: walls, periodic faces, anisotropic spacing)."""
    z = load_golden(os.path.basename(path))
    spec = _blockspec(z)
    print(f"case {os.path.basename(path)} {smoother} {spec}", flush=True)
    with _solver(spec, use_graph=False, smoother=smoother) as S:
        _relax_and_vcycle(S, z, scheme, spec["nlev"])


@pytest.mark.parametrize("smoother,scheme", SMOOTHERS, ids=[s for s, _ in SMOOTHERS])
@pytest.mark.parametrize("path", LSM_LINE, ids=[os.path.basename(p) for p in LSM_LINE])
def test_gpu_line_smoothers_variable_density_bitwise(gpu_lib, path, smoother, scheme):
    """The same under L_LSM: the per-cell density-weighted coefficients feed
    the GS update and the Thomas recurrences; bitwise the reference build."""
    z = load_golden(os.path.basename(path))
    spec = _blockspec(z)
    print(f"case {os.path.basename(path)} {smoother} {spec}", flush=True)
    with _solver(spec, use_graph=False, smoother=smoother) as S:
        S.set_density(0, z["dens"])
        _relax_and_vcycle(S, z, scheme, spec["nlev"])


@pytest.mark.parametrize("graph", [True, False], ids=["graph", "eager"])
@pytest.mark.parametrize("smoother,scheme", SMOOTHERS, ids=[s for s, _ in SMOOTHERS])
@pytest.mark.parametrize("group", ["whole_sweeps", "wide_sweeps", "diagonals"])
def test_gpu_line_smoothers_solve_bitwise_vs_oracle(gpu_lib, smoother, scheme, graph, group):
    """Unequal blocks in one handle (their hyperplanes and line diagonals
    share launches): six cycles of each are bitwise the oracle's solve of
    that block alone.  TDMA levels of up to 4096 cells per diagonal run as
    whole sweeps in one workgroup per block, the previous diagonal's solution
    kept in LDS ("whole_sweeps"; "wide_sweeps": 64 x 64 and 64 x 48 x 64
    diagonals, 130 KB of LDS); a 128 x 40 finest level makes its group run
    diagonal by diagonal over workgroups of lines ("diagonals")."""
    from oracle import oracle as O
    rng = np.random.default_rng(17)
    if group == "whole_sweeps":
        specs = [dict(n=(32, 32, 32), h=(1 / 32,) * 3, bc=(1, 1, 1, 1, 0, 0), nlev=3),
                 dict(n=(24, 16, 40), h=(1 / 24, 1 / 16, 1 / 40), bc=(0, 0, 1, 1, 0, 0), nlev=3),
                 dict(n=(48, 32, 40), h=(1 / 48, 1 / 32, 1 / 40), bc=(1, 1, 0, 0, 1, 1), nlev=3)]
    elif group == "wide_sweeps":
        specs = [dict(n=(64, 64, 32), h=(1 / 64, 1 / 64, 1 / 32), bc=(1, 1, 0, 0, 1, 1), nlev=3),
                 dict(n=(64, 48, 64), h=(1 / 64, 1 / 48, 1 / 64), bc=(0, 0, 1, 1, 0, 0), nlev=3)]
    else:
        specs = [dict(n=(128, 40, 16), h=(1 / 128, 1 / 40, 1 / 16), bc=(1, 1, 0, 0, 1, 1),
                      nlev=3),
                 dict(n=(32, 32, 32), h=(1 / 32,) * 3, bc=(0, 0, 1, 1, 0, 0), nlev=3),
                 # a thin block: its coarsest level is one cell deep (k-lines of one cell)
                 dict(n=(32, 16, 4), h=(1 / 32, 1 / 16, 1 / 4), bc=(0,) * 6, nlev=3)]
    fs = []
    for s in specs:
        f = rng.standard_normal(s["n"][::-1])
        fs.append(f - f.mean())
    nc = 6
    print(f"case {smoother} graph={graph} {group}", flush=True)
    with _solver(specs, maxcy=nc, tol=0.0, use_graph=graph, smoother=smoother) as S:
        for b, f in enumerate(fs):
            S.set_rhs(b, f)
        S.solve()
        got = [S.get_p(b) for b in range(len(specs))]
    for b, (s, f) in enumerate(zip(specs, fs)):
        B = O.Block(s["n"], s["h"], s["bc"], s["nlev"])
        po, _ = O.solve(B, f, nc, 0.0, scheme=scheme)
        assert np.array_equal(got[b][1:-1, 1:-1, 1:-1], po[1:-1, 1:-1, 1:-1]), f"block {b}"


def test_gpu_line_smoothers_refuse_mixed_tail_fused(gpu_lib):
    """GS and TDMA run in FP64 on the level-by-level path: mixed precision,
    the LDS tail and fused red-black sweeps are refused."""
    from hydro3d_amd import _lib as L
    spec = dict(n=(16, 16, 16), h=(1 / 16,) * 3, bc=(1, 1, 1, 1, 0, 0), nlev=3)
    for sm in ("gs", "tdma"):
        with pytest.raises(L.H3dError) as e:
            _solver(spec, smoother=sm, precision="mixed")
        assert e.value.status == L.ERR_UNSUPPORTED
        with _solver(spec, smoother=sm) as S:
            assert S.get_tail() == 4  # no tail level
            with pytest.raises(L.H3dError):
                S.set_tail(2)
            with pytest.raises(L.H3dError):
                S.set_fused(1 << 20)


@pytest.mark.parametrize("smoother,name", [("gs", "newsolv_gs_ppw16.npz"),
                                           ("tdma", "newsolv_tdma_ppw16.npz")])
def test_gpu_line_smoothers_velocity_loop_bitwise_vs_reference(gpu_lib, smoother, name):
    """newsolv_mg's velocity-form loop on the device (the drop-in's path)
    with mg_itrsch 1 and 2: from the reference's velocity field and p = 0,
    the rmax of every iteration, p (ghosts included) and su are bitwise the
    reference build's; with eps = 1e-6 it stops where the reference stops."""
    from hydro3d_amd import _lib as L
    from hydro3d_amd import configs as C
    z = load_golden(name)
    c = C.cube(16, C.BC_PPN, 3, seed=6)
    spec = _blockspec(z)
    nc = int(z["ncycles"])
    print(f"case {name}", flush=True)
    vel = ([_dev(c["u"])], [_dev(c["v"])], [_dev(c["w"])])
    with _solver(spec, maxcy=nc, tol=0.0, dt=float(z["dt"]), smoother=smoother) as S:
        S.set_p(0, np.zeros(c["f"].shape))
        st, ran, hist, _ = S.solve_velocity(*vel)
        p = S.get_p(0)
        su = S.get_su(0).cpu().numpy()
    assert st == L.ERR_NOT_CONVERGED and ran == nc
    assert np.array_equal(hist, z["rmax"]), (hist, z["rmax"])
    assert np.array_equal(p, z["p"])
    assert np.array_equal(su, z["final_su"])
    with _solver(spec, maxcy=nc, tol=1e-6, dt=float(z["dt"]), smoother=smoother) as S:
        S.set_p(0, np.zeros(c["f"].shape))
        st, ran, hist, _ = S.solve_velocity(*vel, raise_on_fail=True)
    assert ran == int(z["stop_at"]) and np.array_equal(hist, z["rmax"][:ran + 1])
