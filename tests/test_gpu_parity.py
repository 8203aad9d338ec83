"""Parity of the HIP path (libh3dmg.so through its C-ABI) with the oracle and
the reference fixtures.  Needs an MI355X.

Bar: bitwise for every FP64 operator and for whole V-cycles and solves
(the library is built with -ffp-contract=off and follows the reference's
arithmetic order); residual norms to 1e-12 relative (reduction order
differs); the mixed-precision path to the tolerances stated per test.
"""
import glob
import os

import numpy as np
import pytest

from h3d_util import (GOLDEN, apply_np, image_ghosts, load_golden, mean_free, rel_l2,
                      spectral_solve)

pytestmark = pytest.mark.gpu

OPS = sorted(glob.glob(os.path.join(GOLDEN, "ops_*.npz")))


def _blockspec(z):
    return dict(n=tuple(int(x) for x in z["n"]), h=tuple(float(x) for x in z["h"]),
                bc=tuple(int(x) for x in z["bc"]), nlev=int(z["nlev"]))


def _dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).cuda()


def _host(t):
    return t.cpu().numpy()


def _solver(specs, **kw):
    from hydro3d_amd import MultigridSolver
    return MultigridSolver(specs if isinstance(specs, list) else [specs], **kw)


@pytest.mark.parametrize("path", OPS, ids=[os.path.basename(p) for p in OPS])
def test_gpu_operators_bitwise_vs_reference(gpu_lib, path):
    import torch
    from hydro3d_amd import _lib as L
    z = load_golden(os.path.basename(path))
    spec = _blockspec(z)
    Lv = spec["nlev"]
    print(f"case {os.path.basename(path)} {spec}", flush=True)
    with _solver(spec, use_graph=False) as S:
        for g in range(1, Lv + 1):
            phi, rhs = z[f"relax{g}_s3_in_phi"], z[f"relax{g}_s3_in_rhs"]
            S.dbg_put(0, g, L.ARR_PHI, _dev(phi))
            S.dbg_put(0, g, L.ARR_RHS, _dev(rhs))
            S.dbg_relax(g, 2)
            out = _host(S.dbg_get(0, g, L.ARR_PHI, torch.empty(phi.shape, dtype=torch.float64,
                                                                 device="cuda")))
            assert np.array_equal(out, z[f"relax{g}_s3_out"]), f"relax level {g}"
            if g > 1:
                S.dbg_put(0, g - 1, L.ARR_PHI, _dev(z[f"restr{g}_in_phi"]))
                S.dbg_put(0, g - 1, L.ARR_RHS, _dev(z[f"restr{g}_in_rhs"]))
                S.dbg_restrict(g)
                pc = _host(S.dbg_get(0, g, L.ARR_PHI, torch.empty(z[f"restr{g}_out_phi"].shape,
                                                                    dtype=torch.float64,
                                                                    device="cuda")))
                rc = _host(S.dbg_get(0, g, L.ARR_RHS, torch.empty(z[f"restr{g}_out_rhs"].shape,
                                                                    dtype=torch.float64,
                                                                    device="cuda")))
                assert np.array_equal(pc, z[f"restr{g}_out_phi"]), f"restr phi level {g}"
                assert np.array_equal(rc, z[f"restr{g}_out_rhs"]), f"restr rhs level {g}"
            if g < Lv:
                S.dbg_put(0, g, L.ARR_PHI, _dev(z[f"corr{g}_in_fine"]))
                S.dbg_put(0, g + 1, L.ARR_PHI, _dev(z[f"corr{g}_in_coarse"]))
                S.dbg_prolong(g)
                out = _host(S.dbg_get(0, g, L.ARR_PHI, torch.empty(z[f"corr{g}_out"].shape,
                                                                     dtype=torch.float64,
                                                                     device="cuda")))
                assert np.array_equal(out, z[f"corr{g}_out"]), f"corr level {g}"
        S.set_rhs(0, z["vcycle_s3_su"])
        S.vcycle()
        pp = S.get_pp(0)
        assert np.array_equal(pp, z["vcycle_s3_pp"]), "vcycle"


LSM_OPS = sorted(glob.glob(os.path.join(GOLDEN, "lsm_ops_*.npz")))


@pytest.mark.parametrize("path", LSM_OPS, ids=[os.path.basename(p) for p in LSM_OPS])
def test_gpu_lsm_operators_bitwise_vs_reference(gpu_lib, path):
    """Variable-density (L_LSM) operator: from the reference build's density,
    the per-level coefficients (coeff's LSM branch, dens_mg restricted level
    by level) under relax, restriction and a whole V-cycle are bitwise the
    reference's."""
    import torch
    from hydro3d_amd import _lib as L
    z = load_golden(os.path.basename(path))
    spec = _blockspec(z)
    Lv = spec["nlev"]
    print(f"case {os.path.basename(path)} {spec}", flush=True)
    with _solver(spec, use_graph=False) as S:
        S.set_density(0, z["dens"])
        for g in range(1, Lv + 1):
            phi, rhs = z[f"relax{g}_in_phi"], z[f"relax{g}_in_rhs"]
            S.dbg_put(0, g, L.ARR_PHI, _dev(phi))
            S.dbg_put(0, g, L.ARR_RHS, _dev(rhs))
            S.dbg_relax(g, 2)
            out = _host(S.dbg_get(0, g, L.ARR_PHI, torch.empty(phi.shape, dtype=torch.float64,
                                                                 device="cuda")))
            assert np.array_equal(out, z[f"relax{g}_out"]), f"relax level {g}"
            if g > 1:
                S.dbg_put(0, g - 1, L.ARR_PHI, _dev(z[f"restr{g}_in_phi"]))
                S.dbg_put(0, g - 1, L.ARR_RHS, _dev(z[f"restr{g}_in_rhs"]))
                S.dbg_restrict(g)
                rc = _host(S.dbg_get(0, g, L.ARR_RHS, torch.empty(z[f"restr{g}_out_rhs"].shape,
                                                                    dtype=torch.float64,
                                                                    device="cuda")))
                assert np.array_equal(rc, z[f"restr{g}_out_rhs"]), f"restr rhs level {g}"
        S.set_rhs(0, z["vcycle_su"])
        S.vcycle()
        assert np.array_equal(S.get_pp(0), z["vcycle_pp"]), "vcycle"


@pytest.mark.parametrize("name,smoother", [("lsm_newsolv_ppw16.npz", "rbgs"),
                                           ("lsm_newsolv_tdma_ppw16.npz", "tdma")])
def test_gpu_lsm_velocity_loop_bitwise_vs_reference(gpu_lib, name, smoother):
    """The velocity-form loop under L_LSM (calvel over the face density,
    calmas writing sup, the density-weighted V-cycle), with red-black and
    with line TDMA (the SolitaryWave benchmark's combination): rmax of every
    iteration, the final pressure and the final sup are bitwise the reference
    build's."""
    from hydro3d_amd import configs as C
    from hydro3d_amd import _lib as L
    z = load_golden(name)
    c = C.cube(16, C.BC_PPN, 3, seed=5)
    spec = _blockspec(z)
    nc = int(z["ncycles"])
    print(f"case {name}", flush=True)
    with _solver(spec, maxcy=nc, tol=0.0, dt=float(z["dt"]), smoother=smoother) as S:
        S.set_density(0, z["dens"])
        S.set_p(0, np.zeros(c["f"].shape))
        st, ran, hist, _ = S.solve_velocity([_dev(c["u"])], [_dev(c["v"])], [_dev(c["w"])])
        p = S.get_p(0)
        sup = _host(S.get_su(0))
    assert st == L.ERR_NOT_CONVERGED and ran == nc
    assert np.array_equal(hist, z["rmax"]), (hist, z["rmax"])
    assert np.array_equal(p, z["p"])
    assert np.array_equal(sup, z["final_sup"])


def test_gpu_lsm_refuses_mixed_and_tail(gpu_lib):
    from hydro3d_amd import _lib as L
    spec = dict(n=(16, 16, 16), h=(1 / 16,) * 3, bc=(1, 1, 1, 1, 0, 0), nlev=3)
    dens = np.ones((16, 16, 16))
    with _solver(spec, precision="mixed", coarse="exact") as S:
        with pytest.raises(L.H3dError) as e:
            S.set_density(0, dens)
        assert e.value.status == L.ERR_UNSUPPORTED
    with _solver(spec) as S:
        S.set_density(0, dens)
        assert S.get_tail() == S.nlevels_max + 1
        with pytest.raises(L.H3dError):
            S.set_tail(1)
        with pytest.raises(L.H3dError):
            S.set_fused(1 << 20)


def test_gpu_solve_tracks_reference_velocity_form(gpu_lib):
    from oracle import oracle as O
    z = load_golden("newsolv_ppw32.npz")
    spec = _blockspec(z)
    nc = int(z["ncycles"])
    f = z["su0"]
    with _solver(spec, maxcy=nc, tol=0.0) as S:
        S.set_rhs(0, f)
        st, ran, hist = S.solve()
        assert ran == nc
        p = S.get_p(0)
    fn = float(z["su_norm"][0])
    # reference history (velocity form) per cycle
    assert np.all(np.abs(hist * fn - z["su_norm"]) <= 1e-10 * fn)
    assert rel_l2(mean_free(p[1:-1, 1:-1, 1:-1]), mean_free(z["p"][1:-1, 1:-1, 1:-1])) <= 1e-10
    # and bitwise the oracle's pressure form
    B = O.Block(spec["n"], spec["h"], spec["bc"], spec["nlev"])
    po, ho = O.solve(B, f, nc, 0.0)
    assert np.array_equal(p[1:-1, 1:-1, 1:-1], po[1:-1, 1:-1, 1:-1])
    assert np.allclose(hist * fn, ho, rtol=1e-12, atol=0)


@pytest.mark.parametrize("name", ["newsolv_ppw32.npz", "newsolv_ppw64.npz"])
def test_gpu_reference_stopping_rule(gpu_lib, name):
    """H3DMG_CRIT_REF_RMAX is the reference's rule (rmax < eps and iter > 1,
    newsolv_mg.for:57, rmax = max|div u| dx dy dz from calmas): from the
    reference's su0 the device stops at the cycle the reference stops at, and
    its rmax per cycle is the reference's to 1e-10 relative."""
    from hydro3d_amd import configs as C
    z = load_golden(name)
    spec = _blockspec(z)
    if "su0" in z:
        f = z["su0"]
    else:  # rebuild the 64^3 workload from its recipe (no mean removal: su0)
        f = C.divergence_rhs(*C.face_velocity(spec["n"], spec["bc"], 0, 0.0, spec["h"]),
                             spec["h"], float(z["dt"]), remove_mean=False)
    eps = 1e-6  # the eps the fixtures were made with
    stop = int(z["stop_at"])
    with _solver(spec, maxcy=int(z["ncycles"]), tol=eps, criterion="ref_rmax",
                 dt=float(z["dt"]), min_cycles=2) as S:
        S.set_rhs(0, f)
        st, ran, hist = S.solve(raise_on_fail=True)
    assert ran == stop, (ran, stop)
    ref = z["rmax"][:stop + 1]
    assert np.all(np.abs(hist - ref) <= 1e-10 * ref), (hist, ref)


def _velocity_case(name):
    from hydro3d_amd import configs as C
    return C.cube(32, C.BC_PPN, 3, seed=0) if name == "newsolv_ppw32.npz" else C.single_64()


@pytest.mark.parametrize("graph", [True, False], ids=["graph", "eager"])
@pytest.mark.parametrize("name", ["newsolv_ppw32.npz", "newsolv_ppw64.npz"])
def test_gpu_velocity_loop_bitwise_vs_reference(gpu_lib, name, graph):
    """h3dmg_solve_velocity is newsolv_mg's loop on the device (calvel,
    calmas, then per iteration mgkcyc from su, p = p + pp, calvel, calmas):
    from the reference's velocity field and p = 0, the rmax of every iteration,
    the final pressure (ghosts included) and the final su are bitwise the
    reference build's; with eps = 1e-6 it stops where the reference stops."""
    from hydro3d_amd import _lib as L
    z = load_golden(name)
    c = _velocity_case(name)
    spec = _blockspec(z)
    nc = int(z["ncycles"])
    print(f"case {name} graph={graph}", flush=True)
    vel = ([_dev(c["u"])], [_dev(c["v"])], [_dev(c["w"])])
    with _solver(spec, maxcy=nc, tol=0.0, dt=float(z["dt"]), use_graph=graph) as S:
        S.set_p(0, np.zeros(c["f"].shape))
        st, ran, hist, _ = S.solve_velocity(*vel)
        p = S.get_p(0)
        su = _host(S.get_su(0))
    assert st == L.ERR_NOT_CONVERGED and ran == nc
    assert np.array_equal(hist, z["rmax"]), (hist, z["rmax"])
    pin = p[1:-1, 1:-1, 1:-1]
    if "p" in z:
        assert np.array_equal(p, z["p"])
        assert np.array_equal(su, z["final_su"])
    # scalar summaries (the 64^3 fixture keeps no arrays): numpy's sum and
    # norm depend on the array's alignment in their last bits
    assert abs(pin.sum() - float(z["p_sum"])) <= 1e-14 * abs(float(z["p_sum"])) + 1e-300
    p_l2 = np.linalg.norm(pin - pin.mean())
    assert abs(p_l2 - float(z["p_l2_meanfree"])) <= 1e-13 * float(z["p_l2_meanfree"])
    assert abs(np.linalg.norm(su) - float(z["final_su_l2"])) <= 1e-14 * float(z["final_su_l2"])
    with _solver(spec, maxcy=nc, tol=1e-6, dt=float(z["dt"]), use_graph=graph) as S:
        S.set_p(0, np.zeros(c["f"].shape))
        st, ran, hist, _ = S.solve_velocity(*vel, raise_on_fail=True)
    assert ran == int(z["stop_at"]) and np.array_equal(hist, z["rmax"][:ran + 1])


def test_gpu_velocity_loop_from_host_arrays(gpu_lib):
    """The Fortran-layout entry points (h3dmg_set_velocity, h3dmg_get_velocity,
    h3dmg_get_su: two ghost layers, u(i) the face between cells i and i+1)
    feed the same loop: from the reference's velocity field the rmax history,
    p and su are bitwise the reference's, and u, v, w land on the faces they
    came from, bitwise the device-array path's; faces calvel does not
    correct keep the host's u."""
    z = load_golden("newsolv_ppw32.npz")
    c = _velocity_case("newsolv_ppw32.npz")
    spec = _blockspec(z)
    ni, nj, nk = spec["n"]
    nc, pl = int(z["ncycles"]), 2
    shape = (nk + 2 * pl, nj + 2 * pl, ni + 2 * pl)
    rng = np.random.default_rng(7)
    US, VS, WS = (rng.standard_normal(shape) for _ in range(3))  # ghosts: noise
    U, V, W = (rng.standard_normal(shape) for _ in range(3))
    US[pl:pl + nk, pl:pl + nj, pl - 1:pl + ni] = c["u"]
    VS[pl:pl + nk, pl - 1:pl + nj, pl:pl + ni] = c["v"]
    WS[pl - 1:pl + nk, pl:pl + nj, pl:pl + ni] = c["w"]
    # bc (P, P, P, P, W, W): only the two z-wall faces are not corrected;
    # the host's u there must be what calvel leaves (the reference keeps u)
    W[pl - 1, pl:pl + nj, pl:pl + ni] = c["w"][0]
    W[pl + nk - 1, pl:pl + nj, pl:pl + ni] = c["w"][-1]
    with _solver(spec, maxcy=nc, tol=0.0, dt=float(z["dt"])) as S:
        S.set_p(0, np.zeros(c["f"].shape))
        S.set_velocity(0, US, VS, WS, U, V, W, pl=pl)
        st, ran, hist, _ = S.solve_velocity()
        p = S.get_p(0)
        su = S.get_su_host(0, pl=pl)
        uo, vo, wo = S.get_velocity(0, pl=pl)
    assert ran == nc and np.array_equal(hist, z["rmax"])
    assert np.array_equal(p, z["p"])
    assert np.array_equal(su[pl:pl + nk, pl:pl + nj, pl:pl + ni], z["final_su"])
    with _solver(spec, maxcy=nc, tol=0.0, dt=float(z["dt"])) as S:
        S.set_p(0, np.zeros(c["f"].shape))
        _, _, _, (ud, vd, wd) = S.solve_velocity([_dev(c["u"])], [_dev(c["v"])], [_dev(c["w"])])
    assert np.array_equal(uo[pl:pl + nk, pl:pl + nj, pl - 1:pl + ni], _host(ud[0]))
    assert np.array_equal(vo[pl:pl + nk, pl - 1:pl + nj, pl:pl + ni], _host(vd[0]))
    assert np.array_equal(wo[pl - 1:pl + nk, pl:pl + nj, pl:pl + ni], _host(wd[0]))


def test_gpu_velocity_loop_mixed(gpu_lib):
    """The velocity form with the FP32 V-cycle and exact coarsest solve:
    each iteration's su is recomputed in FP64 from the corrected velocity, so
    the loop converges past what the FP32 cycle alone resolves; the final
    pressure is within 1e-7 of the FP64 loop's converged pressure."""
    c = _velocity_case("newsolv_ppw64.npz")
    spec = dict(n=c["n"], h=c["h"], bc=c["bc"], nlev=c["nlev"])
    vel = ([_dev(c["u"])], [_dev(c["v"])], [_dev(c["w"])])
    out = {}
    for prec, tol in (("fp64", 1e-13), ("mixed", 1e-11)):
        kw = dict(precision="mixed", coarse="exact") if prec == "mixed" else {}
        with _solver(spec, maxcy=60, tol=tol, dt=c["dt"], **kw) as S:
            S.set_p(0, np.zeros(c["f"].shape))
            st, ran, hist, _ = S.solve_velocity(*vel, raise_on_fail=True)
            out[prec] = (ran, hist, mean_free(S.get_p(0)[1:-1, 1:-1, 1:-1]))
        print(prec, out[prec][0], out[prec][1][-1], flush=True)
    assert out["mixed"][1][-1] < 1e-11
    assert rel_l2(out["mixed"][2], out["fp64"][2]) <= 1e-7


def test_gpu_single_64_to_tolerance(gpu_lib):
    """64^3, periodic x/y, Neumann z, 4 levels, V(2,2), to 1e-8: same cycle
    count and bitwise the same pressure as the oracle."""
    from hydro3d_amd import configs as C
    from oracle import oracle as O
    c = C.single_64()
    spec = dict(n=c["n"], h=c["h"], bc=c["bc"], nlev=c["nlev"])
    with _solver(spec, tol=1e-8) as S:
        S.set_rhs(0, c["f"])
        st, ran, hist = S.solve(raise_on_fail=True)
        p = S.get_p(0)
    B = O.Block(c["n"], c["h"], c["bc"], c["nlev"])
    po, ho = O.solve(B, c["f"], 60, 1e-8)
    assert ran == len(ho) - 1
    assert np.array_equal(p[1:-1, 1:-1, 1:-1], po[1:-1, 1:-1, 1:-1])
    assert hist[-1] <= 1e-8


def test_gpu_uneven_blocks_each_equal_alone(gpu_lib):
    """Independent blocks of different shapes and level counts in one handle:
    each block's pressure is bitwise the oracle's single-block solve."""
    from oracle import oracle as O
    rng = np.random.default_rng(3)
    W, P = 0, 1
    # level counts 4, 3, 3: level 3 is the coarsest of two blocks and an
    # intermediate level of the first (10, 17 and 21 cycles in the oracle)
    specs = [dict(n=(32, 16, 24), h=(1 / 32,) * 3, bc=(W,) * 6, nlev=4),
             dict(n=(16, 16, 16), h=(1 / 16,) * 3, bc=(P, P, P, P, W, W), nlev=3),
             dict(n=(48, 8, 16), h=(0.02, 0.03, 0.04), bc=(P, P, W, W, P, P), nlev=3)]
    fs = [mean_free(rng.standard_normal((s["n"][2], s["n"][1], s["n"][0]))) for s in specs]
    with _solver(specs, tol=1e-9) as S:
        for b, f in enumerate(fs):
            S.set_rhs(b, f)
        st, ran, hist = S.solve(raise_on_fail=True)
        cyc, rel = S.block_stats()
        ps = [S.get_p(b) for b in range(len(specs))]
    for b, (s, f) in enumerate(zip(specs, fs)):
        B = O.Block(s["n"], s["h"], s["bc"], s["nlev"])
        po, ho = O.solve(B, f, 60, 1e-9)
        assert cyc[b] == len(ho) - 1, (b, cyc[b], len(ho) - 1)
        assert np.array_equal(ps[b][1:-1, 1:-1, 1:-1], po[1:-1, 1:-1, 1:-1]), b


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_gpu_tail_level_does_not_change_results(gpu_lib, prec):
    """The single-workgroup V-cycle tail computes every cell with the same
    arithmetic as the per-level kernels: the tail level (auto, every level,
    off) changes neither the pressure (bitwise) nor the cycle counts."""
    rng = np.random.default_rng(5)
    W, P = 0, 1
    specs = [dict(n=(32, 16, 24), h=(1 / 32,) * 3, bc=(W,) * 6, nlev=4),
             dict(n=(16, 16, 16), h=(1 / 16,) * 3, bc=(P, P, P, P, W, W), nlev=3),
             dict(n=(48, 8, 16), h=(0.02, 0.03, 0.04), bc=(P, P, W, W, P, P), nlev=3)]
    fs = [mean_free(rng.standard_normal((s["n"][2], s["n"][1], s["n"][0]))) for s in specs]
    kw = dict(tol=1e-9) if prec == "fp64" else dict(tol=1e-9, precision="mixed", coarse="exact")
    from hydro3d_amd import _lib as L
    results = {}
    for tail in (0, 1, 2, 3, 4, 5):
        print(f"case: tail level {tail} ({prec})", flush=True)
        with _solver(specs, **kw) as S:
            try:
                S.set_tail(tail)
            except L.H3dError as e:  # the levels from `tail` down do not fit in LDS
                assert e.status == L.ERR_ARG and tail > 0
                continue
            used = S.get_tail()
            for b, f in enumerate(fs):
                S.set_rhs(b, f)
            S.solve(raise_on_fail=True)
            cyc, _ = S.block_stats()
            results[tail] = (used, list(cyc), [S.get_p(b) for b in range(len(specs))])
    # auto: with the exact solve the shallowest coarsest level (3 here), off
    # with a smoothed coarsest level; 5 is above every coarsest level: off
    assert results[0][0] == (3 if prec == "mixed" else 5) and results[5][0] == 5
    assert {2, 3, 4} <= set(results), sorted(results)
    ref = results[5]
    for tail, (used, cyc, ps) in results.items():
        assert cyc == ref[1], (tail, cyc, ref[1])
        for b in range(len(specs)):
            assert np.array_equal(ps[b], ref[2][b]), (tail, b)


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_gpu_march_depth_does_not_change_results(gpu_lib, prec):
    """Smoother and residual tiles of any depth (1..16 planes, including the
    software-pipelined 8-plane interior tiles) compute every cell with the
    same arithmetic: the pressure is bitwise that of the automatic choice."""
    rng = np.random.default_rng(11)
    W, P = 0, 1
    specs = [dict(n=(64, 64, 64), h=(1 / 64,) * 3, bc=(P, P, P, P, W, W), nlev=4),
             dict(n=(32, 48, 40), h=(0.02, 0.03, 0.025), bc=(W,) * 6, nlev=3)]
    fs = [mean_free(rng.standard_normal((s["n"][2], s["n"][1], s["n"][0]))) for s in specs]
    kw = dict(tol=1e-9, maxcy=12)
    if prec == "mixed":
        kw.update(precision="mixed", coarse="exact")
    results = {}
    for depth in (0, 1, 2, 4, 8, 16):
        print(f"case: march depth {depth} ({prec})", flush=True)
        with _solver(specs, **kw) as S:
            S.set_tail(S.nlevels_max + 1)  # every level on the per-level kernels
            for g in range(1, S.nlevels_max + 1):
                S.set_march(g, depth)
            for b, f in enumerate(fs):
                S.set_rhs(b, f)
            S.solve()  # 12 cycles whether or not they reach 1e-9
            cyc, _ = S.block_stats()
            results[depth] = (list(cyc), [S.get_p(b) for b in range(len(specs))])
    ref = results[0]
    for depth, (cyc, ps) in results.items():
        assert cyc == ref[0], (depth, cyc, ref[0])
        for b in range(len(specs)):
            assert np.array_equal(ps[b], ref[1][b]), (depth, b)


@pytest.mark.parametrize("n", [(64, 64, 64), (128, 128, 128), (128, 128, 256), (128, 256, 256)],
                         ids=["64", "128", "128x128x256", "128x256x256"])
def test_gpu_residual_norm_matches_numpy(gpu_lib, n):
    """The fine-level residual kernel at shapes whose tiles are 1, 2, 4 and 8
    planes deep, boundary tiles included: the relative
    residual norm of a random pressure, hist[0] of a solve, equals the numpy
    norm of f - A p to 1e-12 (only the reduction order differs)."""
    rng = np.random.default_rng(3)
    P, W = 1, 0
    bc = (P, P, P, P, W, W)
    h = tuple(1.0 / x for x in n)
    spec = dict(n=n, h=h, bc=bc, nlev=3)
    f = mean_free(rng.standard_normal((n[2], n[1], n[0])))
    p = rng.standard_normal((n[2], n[1], n[0]))
    pad = np.zeros((n[2] + 2, n[1] + 2, n[0] + 2))
    pad[1:-1, 1:-1, 1:-1] = p
    r = f - apply_np(image_ghosts(pad, bc), h, bc)
    want = float(np.linalg.norm(r) / np.linalg.norm(f))
    for prec in ("fp64", "mixed"):
        print(f"case: residual {n} {prec}", flush=True)
        kw = dict(precision="mixed", coarse="exact") if prec == "mixed" else {}
        with _solver(spec, tol=1e-30, maxcy=1, **kw) as S:
            S.set_rhs(0, f)
            S.set_p(0, p)
            _, _, hist = S.solve()
        assert abs(hist[0] - want) <= 1e-12 * want, (prec, hist[0], want)


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_gpu_fused_sweeps_do_not_change_results(gpu_lib, prec):
    """A fused red-black sweep (one launch per sweep, red halo recomputed per
    tile) computes every cell as the two split half-sweeps do: with fusion
    off, on the default levels and on every level, the pressure is bitwise
    the same (odd, ragged and one-tile-wide shapes included)."""
    rng = np.random.default_rng(13)
    W, P = 0, 1
    specs = [dict(n=(64, 64, 64), h=(1 / 64,) * 3, bc=(P, P, P, P, W, W), nlev=4),
             dict(n=(48, 40, 24), h=(0.02, 0.03, 0.04), bc=(W,) * 6, nlev=3),
             dict(n=(40, 8, 12), h=(0.025, 0.125, 0.1), bc=(P, P, W, W, P, P), nlev=2)]
    fs = [mean_free(rng.standard_normal((s["n"][2], s["n"][1], s["n"][0]))) for s in specs]
    kw = dict(tol=1e-9, maxcy=10)
    if prec == "mixed":
        kw.update(precision="mixed", coarse="exact")
    results = {}
    for fused in (0, 1 << 18, 1 << 40):
        print(f"case: fused threshold {fused} ({prec})", flush=True)
        with _solver(specs, **kw) as S:
            S.set_tail(S.nlevels_max + 1)  # every level on the per-level kernels
            S.set_fused(fused)
            assert S.get_fused() == fused
            for b, f in enumerate(fs):
                S.set_rhs(b, f)
            S.solve()
            cyc, _ = S.block_stats()
            results[fused] = (list(cyc), [S.get_p(b) for b in range(len(specs))])
    ref = results[0]
    for fused, (cyc, ps) in results.items():
        assert cyc == ref[0], (fused, cyc, ref[0])
        for b in range(len(specs)):
            assert np.array_equal(ps[b], ref[1][b]), (fused, b)


def test_gpu_projection_bitwise_vs_reference(gpu_lib):
    """calvel + calmas on the device (h3dmg_project_device) reproduce the
    reference's su bitwise: su0 from p = 0 and the final su from the
    reference's final pressure (newsolv_ppw32 fixture), and its rmax."""
    import torch
    from hydro3d_amd import configs as C
    z = load_golden("newsolv_ppw32.npz")
    c = C.cube(32, C.BC_PPN, 3, seed=0)
    spec = dict(n=c["n"], h=c["h"], bc=c["bc"], nlev=c["nlev"])
    us, vs, ws = _dev(c["u"]), _dev(c["v"]), _dev(c["w"])
    with _solver(spec, dt=float(z["dt"])) as S:
        S.set_p(0, np.zeros(c["f"].shape))
        _, _, _, su0, rmax0 = S.project(0, us, vs, ws)
        S.set_p(0, z["p"])
        u, v, w, su, rmax = S.project(0, us, vs, ws)
    assert np.array_equal(_host(su0), z["su0"])
    assert np.array_equal(_host(su), z["final_su"])
    # the driver runs all ncycles cycles: the final p is the last cycle's
    assert rmax == z["rmax"][-1], (rmax, z["rmax"])
    assert rmax0 == z["rmax"][0]
    # periodic x: the wrap face's two copies agree; walls in z keep w* = 0
    assert torch.equal(u[:, :, 0], u[:, :, -1])
    assert float(w[0].abs().max()) == 0.0 and float(w[-1].abs().max()) == 0.0


def test_gpu_block_beyond_level_span_is_refused(gpu_lib):
    """Kernels address a level's arrays as 32-bit offsets: a block whose
    level-1 arrays span 4 GiB or more is refused before any allocation."""
    from hydro3d_amd import _lib as L
    n = (1024, 1024, 1024)
    with pytest.raises(L.H3dError) as e:
        _solver(dict(n=n, h=(1 / 1024,) * 3, bc=(1, 1, 1, 1, 0, 0), nlev=6))
    assert e.value.status == L.ERR_UNSUPPORTED, e.value


def test_gpu_graph_and_eager_identical(gpu_lib):
    from hydro3d_amd import configs as C
    c = C.cube(32, C.BC_PPN, 3, seed=9)
    spec = dict(n=c["n"], h=c["h"], bc=c["bc"], nlev=c["nlev"])
    out = []
    for graph in (True, False, True):
        with _solver(spec, tol=1e-10, use_graph=graph) as S:
            S.set_rhs(0, c["f"])
            S.solve(raise_on_fail=True)
            out.append(S.get_p(0))
    assert np.array_equal(out[0], out[1]) and np.array_equal(out[0], out[2])


@pytest.mark.parametrize("coarse", ["smooth", "exact"])
def test_gpu_mixed_precision_64(gpu_lib, coarse):
    """FP32 V-cycle inside FP64 refinement reaches 1e-8 (checked in numpy).
    Pressure: with the reference's coarsest-level smoothing the mixed path
    follows the FP64 iterates to 1e-7; with the exact coarsest solve it lands
    within 1e-7 of the converged discrete solution (spectral solve), which the
    FP64 reference schedule at 1e-8 does not (about 1.4e-5 here)."""
    from hydro3d_amd import configs as C
    c = C.single_64()
    spec = dict(n=c["n"], h=c["h"], bc=c["bc"], nlev=c["nlev"])
    with _solver(spec, tol=1e-8) as S:
        S.set_rhs(0, c["f"])
        S.solve(raise_on_fail=True)
        p64 = mean_free(S.get_p(0)[1:-1, 1:-1, 1:-1])
    with _solver(spec, tol=1e-8, precision="mixed", coarse=coarse) as S:
        S.set_rhs(0, c["f"])
        st, ran, hist = S.solve(raise_on_fail=True)
        pm_pad = S.get_p(0)
    pm = mean_free(pm_pad[1:-1, 1:-1, 1:-1])
    pstar = spectral_solve(c["f"], c["h"], c["bc"])
    r = c["f"] - apply_np(image_ghosts(pm_pad, c["bc"]), c["h"], c["bc"])
    rel = float(np.linalg.norm(r) / np.linalg.norm(c["f"]))
    e_m, e_64 = rel_l2(pm, pstar), rel_l2(p64, pstar)
    e_m64 = rel_l2(pm, p64)
    print(f"mixed/{coarse}: {ran} cycles, residual {rel:.3e}; pressure error vs converged: "
          f"mixed {e_m:.3e}, fp64@1e-8 {e_64:.3e}; mixed vs fp64 {e_m64:.3e}")
    assert rel <= 1e-8 * 1.0001
    if coarse == "smooth":
        assert e_m64 <= 1e-7
    else:
        assert e_m <= 1e-7


def test_gpu_256_fp64_and_mixed(gpu_lib):
    """256^3 (6 levels): FP64 and mixed both reach 1e-8 (checked in numpy);
    the mixed pressure is within 1e-7 of the converged FP64 solution."""
    from hydro3d_amd import configs as C
    c = C.single_256()
    spec = dict(n=c["n"], h=c["h"], bc=c["bc"], nlev=c["nlev"])
    fn = np.linalg.norm(c["f"])
    res = {}
    for prec, coarse, tol in (("fp64", "smooth", 1e-12), ("fp64", "smooth", 1e-8),
                              ("mixed", "exact", 1e-8)):
        with _solver(spec, tol=tol, precision=prec, coarse=coarse, maxcy=200) as S:
            S.set_rhs(0, c["f"])
            st, ran, hist = S.solve(raise_on_fail=True)
            p = S.get_p(0)
        r = c["f"] - apply_np(p, c["h"], c["bc"])
        rel = float(np.linalg.norm(r) / fn)
        print(f"256^3 {prec}/{coarse} tol {tol:g}: {ran} cycles, device {hist[-1]:.3e}, "
              f"numpy {rel:.3e}", flush=True)
        assert rel <= tol * 1.0001
        res[(prec, tol)] = mean_free(p[1:-1, 1:-1, 1:-1])
        del p, r
    pstar = res[("fp64", 1e-12)]
    e_m = rel_l2(res[("mixed", 1e-8)], pstar)
    e_64 = rel_l2(res[("fp64", 1e-8)], pstar)
    print(f"pressure error vs converged FP64: mixed {e_m:.3e}, fp64@1e-8 {e_64:.3e}")
    assert e_m <= 1e-7


@pytest.mark.parametrize("name", ["newsolv_cds4_ppw32.npz", "newsolv_cds4_wall16.npz"])
def test_gpu_velocity_loop_cds4_bitwise_vs_reference(gpu_lib, name):
    """differencing = 2 (calmas' 4th-order divergence, press.for:17-37): from
    host arrays with two ghost layers -- faces -1 and n+1 included, a wrap
    on periodic axes and seeded host values beyond walls -- the loop's rmax
    at every iteration, the final p (one ghost layer) and su are bitwise the
    reference build's (newsolv_cds4_*.npz)."""
    from hydro3d_amd import _lib as L
    from hydro3d_amd import configs as C
    z = load_golden(name)
    spec = _blockspec(z)
    ni, nj, nk = spec["n"]
    nc, pl = int(z["ncycles"]), 2
    u, v, w = C.face_velocity(spec["n"], spec["bc"], int(z["vel_seed"]), 0.0, spec["h"])
    ue, ve, we = C.extend_faces(u, v, w, spec["bc"], int(z["face_seed"]))
    shape = (nk + 2 * pl, nj + 2 * pl, ni + 2 * pl)
    US, VS, WS = (np.zeros(shape) for _ in range(3))
    US[pl:pl + nk, pl:pl + nj, pl - 2:pl + ni + 1] = ue
    VS[pl:pl + nk, pl - 2:pl + nj + 1, pl:pl + ni] = ve
    WS[pl - 2:pl + nk + 1, pl:pl + nj, pl:pl + ni] = we
    print(f"case {name}", flush=True)
    with _solver(spec, maxcy=nc, tol=0.0, dt=float(z["dt"])) as S:
        S.set_differencing(2)
        S.set_p(0, np.zeros((nk, nj, ni)))
        S.set_velocity(0, US, VS, WS, pl=pl)
        st, ran, hist, _ = S.solve_velocity()
        p = S.get_p(0)
        su = S.get_su_host(0, pl=pl)
    assert st == L.ERR_NOT_CONVERGED and ran == nc
    assert np.array_equal(hist, z["rmax"]), (hist, z["rmax"])
    assert np.array_equal(p, z["p"])
    assert np.array_equal(su[pl:pl + nk, pl:pl + nj, pl:pl + ni], z["final_su"])


def test_gpu_cds4_refusals(gpu_lib):
    """differencing = 2 needs the host arrays' outer faces: device-array
    projections and a loop over device arrays are refused, as is the
    variable-density operator."""
    from hydro3d_amd import _lib as L
    from hydro3d_amd import configs as C
    c = C.cube(16, C.BC_PPN, 3, seed=0)
    spec = dict(n=c["n"], h=c["h"], bc=c["bc"], nlev=c["nlev"])
    with _solver(spec, maxcy=2, tol=0.0, dt=1e-3) as S:
        S.set_differencing(2)
        with pytest.raises(L.H3dError) as e:
            S.project(0, _dev(c["u"]), _dev(c["v"]), _dev(c["w"]))
        assert e.value.status == L.ERR_UNSUPPORTED
        with pytest.raises(L.H3dError) as e:
            S.solve_velocity([_dev(c["u"])], [_dev(c["v"])], [_dev(c["w"])])
        assert e.value.status == L.ERR_UNSUPPORTED
        with pytest.raises(L.H3dError) as e:
            S.set_density(0, np.ones(c["f"].shape))
        assert e.value.status == L.ERR_UNSUPPORTED
        S.set_differencing(1)  # back to the 2nd-order branch: accepted
        st, ran, _, _ = S.solve_velocity([_dev(c["u"])], [_dev(c["v"])], [_dev(c["w"])])
        assert ran == 2
