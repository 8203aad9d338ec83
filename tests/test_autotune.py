"""Autotuner: cache keys and files on the CPU; on the GPU the tuned launch
shapes leave every result bitwise unchanged and the cache round-trips."""
import json
import os

import numpy as np
import pytest

from hydro3d_amd import autotune as AT


class _FakeParams:
    precision, coarse, nu1, nu2, smoother = 1, 1, 2, 2, 3


class _FakeSolver:
    params = _FakeParams()

    def __init__(self, blocks):
        self.blocks = blocks


def test_config_key_ignores_block_order_and_spacing():
    a = dict(n=(32, 16, 24), h=(0.1,) * 3, bc=(0,) * 6, nlev=4)
    b = dict(n=(16, 16, 16), h=(0.2,) * 3, bc=(1, 1, 1, 1, 0, 0), nlev=3)
    k1 = AT.config_key(_FakeSolver([a, b]), device="gfx950")
    k2 = AT.config_key(_FakeSolver([b, dict(a, h=(0.3,) * 3)]), device="gfx950")
    assert k1 == k2
    assert k1 != AT.config_key(_FakeSolver([a]), device="gfx950")
    assert json.loads(k1)["blocks"][0][0] == [16, 16, 16]


def test_line_smoother_handles_have_their_own_key_and_no_knobs():
    """GS / TDMA handles key separately (red-black keys keep their form) and
    the autotuner leaves them alone: their launches have no tunable shape."""
    a = dict(n=(32, 32, 32), h=(0.1,) * 3, bc=(0,) * 6, nlev=3)
    rb = _FakeSolver([a])
    tdma = _FakeSolver([a])
    tdma.params = type("P", (_FakeParams,), {"smoother": 2})()
    k_rb, k_tdma = AT.config_key(rb, device="gfx950"), AT.config_key(tdma, device="gfx950")
    assert "smoother" not in json.loads(k_rb)
    assert json.loads(k_tdma)["smoother"] == 2 and k_rb != k_tdma
    assert AT.autotune(tdma, cache_path=None) == {}


def test_cache_round_trip(tmp_path):
    path = str(tmp_path / "sub" / "tune.json")
    assert AT.load_cache(path) == {}
    AT.save_cache({"k": {"march": {"1": 4}, "tail": 3}}, path)
    assert AT.load_cache(path) == {"k": {"march": {"1": 4}, "tail": 3}}


@pytest.mark.gpu
def test_autotune_keeps_results_and_caches(tmp_path):
    import torch
    from hydro3d_amd import MultigridSolver, configs
    print("case: autotune 64^3 mixed", flush=True)
    c = configs.single_64()
    spec = dict(n=c["n"], h=c["h"], bc=c["bc"], nlev=c["nlev"])
    f = torch.from_numpy(c["f"]).cuda()
    with MultigridSolver([spec], precision="mixed", coarse="exact") as S:
        S.set_rhs(0, f)
        S.solve(raise_on_fail=True)
        p_ref = S.get_p(0)
    path = str(tmp_path / "tune.json")
    with MultigridSolver([spec], precision="mixed", coarse="exact") as S:
        S.set_rhs(0, f)
        choice = AT.autotune(S, path, reps=2)
        S.set_p(0, torch.zeros_like(f))
        S.solve(raise_on_fail=True)
        p_tuned = S.get_p(0)
        assert AT.apply_cached(S, path) == choice
    assert np.array_equal(p_ref, p_tuned)
    assert os.path.exists(path) and len(AT.load_cache(path)) == 1


@pytest.mark.gpu
@pytest.mark.parametrize("prec", ["mixed", "fp64"])
def test_smoother_lanes_per_row_leave_results_bitwise(prec):
    """h3dmg_set_lanes reshapes the smoother's tiles (row segments of fewer
    lanes, more rows per tile): every choice gives the same pressure."""
    import torch
    from hydro3d_amd import MultigridSolver
    spec = dict(n=(128, 96, 64), h=(1 / 64,) * 3, bc=(1, 1, 0, 0, 0, 0), nlev=4)
    rng = np.random.default_rng(21)
    f = rng.standard_normal((64, 96, 128))
    f -= f.mean()
    f = torch.from_numpy(f).cuda()
    kw = dict(precision=prec, coarse="exact" if prec == "mixed" else "smooth", tol=0.0, maxcy=4)
    print(f"case: lanes per row, {prec}", flush=True)
    with MultigridSolver([spec], **kw) as S:
        S.set_rhs(0, f)
        S.solve()
        ref = S.get_p(0, out=torch.empty_like(f))
        natural = [S.get_lanes(g) for g in range(1, 5)]
        assert natural[0] == 32 and S.get_lanes(5) == -1
        for g, lanes in ((1, 16), (1, 4), (2, 8), (2, 1), (3, 2)):
            S.set_lanes(g, lanes)
            assert S.get_lanes(g) == min(lanes, natural[g - 1])
            for kb in (1, 2):
                S.set_march(g, kb)
                S.set_p(0, torch.zeros_like(f))
                S.solve()
                got = S.get_p(0, out=torch.empty_like(f))
                assert torch.equal(got, ref), (g, lanes, kb)
        S.set_lanes(1, 0)
        assert S.get_lanes(1) == natural[0]
