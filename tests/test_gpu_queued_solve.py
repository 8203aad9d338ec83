"""h3dmg_solve in graph mode keeps the next cycle queued and stops on the
device's reading of the loop's rule; eager mode runs the host's loop.  The
two must agree exactly on every way a solve ends: per-block tolerances, the
reference's global rmax criterion (newsolv_mg.for:57), min_cycles, maxcy, a
non-finite right-hand side.  Needs an MI355X."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

H = 1.0 / 32


def _spec(n, bc, nlev):
    return dict(n=tuple(n), h=(H,) * 3, bc=tuple(bc), nlev=nlev)


SPECS = [
    _spec((32, 32, 32), (1, 1, 1, 1, 0, 0), 3),
    _spec((48, 32, 16), (0,) * 6, 3),
    _spec((64, 48, 32), (1, 1, 0, 0, 1, 1), 4),
]


def _rhs(spec, seed, nan=False):
    import torch
    ni, nj, nk = spec["n"]
    g = np.random.default_rng(seed).standard_normal((nk, nj, ni))
    g -= g.mean()
    if nan:
        g[1, 2, 3] = np.nan
    return torch.from_numpy(g).cuda()


def _solve(S):
    """h3dmg_solve's status, cycle count and history, whatever the status."""
    import ctypes
    hist = np.zeros(S.maxcy + 1)
    nc = ctypes.c_int(0)
    fr = ctypes.c_double(0.0)
    st = S.lib.h3dmg_solve(S.h, hist.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), hist.size,
                           ctypes.byref(nc), ctypes.byref(fr))
    return st, nc.value, hist[: nc.value + 1], fr.value


def _run(graph, kw, tols=None, nan_block=None):
    import torch
    from hydro3d_amd import MultigridSolver
    from hydro3d_amd import _lib as L
    with MultigridSolver(SPECS, use_graph=graph, **kw) as S:
        for b, s in enumerate(SPECS):
            S.set_rhs(b, _rhs(s, 40 + b, nan=b == nan_block))
            if tols is not None:
                S.set_block_tol(b, tols[b])
        out = []
        for rep in range(2):  # the second solve starts from the first's pressure
            st, nc, hist, fr = _solve(S)
            assert fr == hist[-1] or not np.isfinite(fr)
            cyc, rel = S.block_stats()
            ps = []
            for b, s in enumerate(SPECS):
                ni, nj, nk = s["n"]
                p = torch.empty((nk, nj, ni), dtype=torch.float64, device="cuda")
                S.get_p(b, out=p)
                ps.append(p)
            out.append((st, nc, hist.copy(), cyc, rel, ps))
            if st == L.ERR_DIVERGED:
                break
    return out


CASES = {
    "rel_l2_per_block_tol": dict(kw=dict(precision="mixed", coarse="exact", tol=1e-8),
                                 tols=[1e-6, 1e-9, 1e-7]),
    "rel_l2_fp64": dict(kw=dict(precision="fp64", coarse="smooth", tol=1e-7)),
    "ref_rmax": dict(kw=dict(precision="fp64", coarse="smooth", criterion="ref_rmax", tol=1e-6,
                             dt=1e-3)),
    "min_cycles": dict(kw=dict(precision="mixed", coarse="exact", tol=1e-2, min_cycles=5)),
    "maxcy": dict(kw=dict(precision="mixed", coarse="exact", tol=1e-14, maxcy=4)),
    "nonfinite": dict(kw=dict(precision="mixed", coarse="exact", tol=1e-8), nan_block=1),
}


@pytest.mark.parametrize("case", sorted(CASES))
def test_gpu_queued_solve_equals_host_loop(gpu_lib, case):
    import torch
    c = CASES[case]
    print(f"case: queued solve vs host loop, {case}", flush=True)
    a = _run(True, c["kw"], c.get("tols"), c.get("nan_block"))
    b = _run(False, c["kw"], c.get("tols"), c.get("nan_block"))
    assert len(a) == len(b)
    for (st1, nc1, h1, cy1, r1, p1), (st2, nc2, h2, cy2, r2, p2) in zip(a, b):
        assert (st1, nc1) == (st2, nc2), (case, st1, nc1, st2, nc2)
        assert cy1 == cy2, (case, cy1, cy2)
        np.testing.assert_array_equal(h1, h2)
        if case != "nonfinite":
            assert r1 == r2, (case, r1, r2)
            for x, y in zip(p1, p2):
                assert torch.equal(x, y), case
