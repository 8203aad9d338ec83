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
