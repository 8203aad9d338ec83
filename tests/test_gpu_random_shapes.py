"""Bitwise parity of one V-cycle with each of the reference's smoothers on
seeded random block shapes, spacings and boundaries, against the oracle.
Needs an MI355X.

The fixtures pin four geometries against the reference build; this sweep
widens the shapes (odd multiples of the level divisor, anisotropic spacing,
mixed periodic and wall axes, several blocks sharing launches) at sizes the
oracle runs in well under a second.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SMOOTHERS = [("rbgs", 3), ("gs", 1), ("tdma", 2)]


def _cases(seed, count):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        nlev = int(rng.integers(2, 4))
        m = 1 << (nlev - 1)
        n = tuple(int(m * rng.integers(1, 48 // m + 1)) for _ in range(3))
        per = [bool(rng.integers(0, 2)) for _ in range(3)]
        bc = tuple(int(per[a // 2]) for a in range(6))
        h = tuple(float(x) for x in rng.uniform(0.5, 2.0, 3) / np.array(n))
        out.append(dict(n=n, h=h, bc=bc, nlev=nlev))
    return out


@pytest.mark.parametrize("smoother,scheme", SMOOTHERS, ids=[s for s, _ in SMOOTHERS])
@pytest.mark.parametrize("seed", [11, 12, 13])
def test_gpu_vcycle_random_shapes_bitwise_vs_oracle(gpu_lib, smoother, scheme, seed):
    """Four random blocks in one handle; each block's V(2,2) cycle from a
    random su equals the oracle's V-cycle of that block alone, bitwise."""
    from hydro3d_amd import MultigridSolver
    from oracle import oracle as O
    specs = _cases(seed, 4)
    rng = np.random.default_rng(seed + 100)
    sus = [rng.standard_normal(s["n"][::-1]) for s in specs]
    print(f"case seed={seed} {smoother} " + " ".join(str(s["n"]) + str(s["bc"][::2]) for s in specs),
          flush=True)
    with MultigridSolver(specs, use_graph=False, smoother=smoother) as S:
        for b, su in enumerate(sus):
            S.set_rhs(b, su)
        S.vcycle()
        got = [S.get_pp(b) for b in range(len(specs))]
    for b, (s, su) in enumerate(zip(specs, sus)):
        B = O.Block(s["n"], s["h"], s["bc"], s["nlev"])
        ref = O.vcycle(B, su, scheme=scheme)
        assert np.array_equal(got[b], ref), f"block {b} {s}"


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_gpu_wide_rows_solve_vs_oracle(gpu_lib, precision):
    """Rows of 128 colour entries or more (ni >= 256): several residual and
    smoother tiles per row.  Blocks 256 wide, 272 wide (a ragged last tile of
    8 entries) and 512 wide, periodic and wall x faces, in one handle: FP64 is bitwise
    the oracle's solve of each block (p, every cycle's residual to 1e-12 --
    the norm's partial sums follow the tiles); the mixed solution's residual
    under the oracle's operator is below 1e-9 of ||f||."""
    from hydro3d_amd import MultigridSolver
    from oracle import oracle as O
    h = (1 / 64,) * 3
    specs = [dict(n=(256, 24, 16), h=h, bc=(1, 1, 0, 0, 0, 0), nlev=3),
             dict(n=(272, 16, 8), h=h, bc=(0, 0, 1, 1, 0, 0), nlev=2),
             dict(n=(512, 8, 8), h=h, bc=(0, 0, 1, 1, 1, 1), nlev=2)]
    rng = np.random.default_rng(21)
    fs = []
    for s in specs:
        f = rng.standard_normal(s["n"][::-1])
        fs.append(f - f.mean())
    nc = 5 if precision == "fp64" else 40
    kw = dict(maxcy=nc, tol=0.0) if precision == "fp64" else dict(
        maxcy=nc, tol=1e-10, precision="mixed", coarse="exact")
    print(f"case wide rows {precision}", flush=True)
    with MultigridSolver(specs, **kw) as S:
        for b, f in enumerate(fs):
            S.set_rhs(b, f)
        _, _, hist = S.solve()
        got = [S.get_p(b) for b in range(len(specs))]
    for b, (s, f) in enumerate(zip(specs, fs)):
        B = O.Block(s["n"], s["h"], s["bc"], s["nlev"])
        if precision == "fp64":
            po, ho = O.solve(B, f, nc, 0.0)
            assert np.array_equal(got[b][1:-1, 1:-1, 1:-1], po[1:-1, 1:-1, 1:-1]), f"block {b}"
        else:
            # the oracle's operator on the device's pressure (ghosts included)
            rel = np.linalg.norm(f - O.apply(B, got[b])) / np.linalg.norm(f)
            assert rel < 1e-9, (b, rel)
    if precision == "fp64":
        # the handle's history is the largest relative residual over blocks
        ref = np.max([O.solve(O.Block(s["n"], s["h"], s["bc"], s["nlev"]), f, nc, 0.0)[1] /
                      np.linalg.norm(f) for s, f in zip(specs, fs)], axis=0)  # relative
        np.testing.assert_allclose(hist, ref, rtol=1e-12)
