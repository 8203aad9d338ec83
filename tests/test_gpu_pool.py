"""Continuous batching on the device: handles built in a pool over block
regions that outlive them, and solves stepped one V-cycle at a time.  Needs
an MI355X.

The bar is bitwise: a block's pressure after its stepped solve -- started
while other blocks were running, carried through handle rebuilds, warm
started from its region's previous pressure -- equals the pressure
h3dmg_solve gives the same block alone.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

H = 1.0 / 64


def _spec(n, bc, nlev):
    return dict(n=tuple(n), h=(H,) * 3, bc=tuple(bc), nlev=nlev)


SPECS = [
    _spec((64, 48, 32), (1, 1, 0, 0, 0, 0), 4),
    _spec((32, 32, 32), (0,) * 6, 3),
    _spec((96, 64, 64), (1,) * 6, 5),
    _spec((48, 80, 16), (0, 0, 1, 1, 0, 0), 3),
]


def _rhs(spec, seed):
    import torch
    ni, nj, nk = spec["n"]
    g = np.random.default_rng(seed).standard_normal((nk, nj, ni))
    g -= g.mean()
    return torch.from_numpy(g).cuda()


def _alone(spec, f, kw, tol, p0=None):
    import torch
    from hydro3d_amd import MultigridSolver
    with MultigridSolver([spec], tol=tol, **kw) as S:
        S.set_rhs(0, f)
        if p0 is not None:
            S.set_p(0, p0)
        st, nc, _ = S.solve()
        cyc, rel = S.block_stats()
        out = torch.empty_like(f)
        S.get_p(0, out=out)
    return out, int(cyc[0]), st


KW = {"mixed": dict(precision="mixed", coarse="exact", maxcy=40, use_graph=False),
      "fp64": dict(precision="fp64", coarse="smooth", maxcy=40, use_graph=True)}


@pytest.mark.parametrize("prec", ["mixed", "fp64"])
def test_gpu_stepped_solve_with_staggered_joins_equals_alone(gpu_lib, prec):
    """Blocks begin at different cycles of one pooled handle; each ends
    with bitwise its solo pressure and cycle count."""
    import torch
    from hydro3d_amd.solver import DevicePool, MultigridSolver, block_bytes
    kw = KW[prec]
    tols = [1e-8, 1e-6, 1e-8, 1e-7]
    fs = [_rhs(s, 100 + q) for q, s in enumerate(SPECS)]
    print(f"case: stepped, staggered joins, {prec}", flush=True)
    with DevicePool(2 << 30, 64) as pool:
        offs = [pool.alloc(block_bytes(s, **kw)) for s in SPECS]
        with MultigridSolver(SPECS, pool=pool, offsets=offs, tol=1e-8, **kw) as S:
            for b, f in enumerate(fs):
                S.set_rhs(b, f)
                S.set_block_tol(b, tols[b])
            S.solve_begin([0, 2])
            steps = 0
            started = {0, 2}
            while True:
                if steps == 2:
                    S.solve_begin([1])
                    started.add(1)
                if steps == 3:
                    S.solve_begin([3])
                    started.add(3)
                n = S.solve_step()
                steps += 1
                if n == 0 and len(started) == 4:
                    break
                assert steps < 200
            got = []
            for b in range(4):
                st = S.block_state(b)
                out = torch.empty_like(fs[b])
                S.get_p(b, out=out)
                got.append((out, st.cycles, st.status, st.running))
        for o in offs:
            pool.free(o)
        assert pool.free_bytes == pool.nbytes
    for b in range(4):
        # (FP64 with the smoothed coarsest level needs more than maxcy cycles
        # on the 3-level blocks: NOT_CONVERGED alone and stepped alike)
        ref, cyc, st = _alone(SPECS[b], fs[b], kw, tols[b])
        assert got[b][1] == cyc and got[b][2] == st and got[b][3] == 0, (b, got[b][1:], cyc, st)
        assert st == 0 or (prec == "fp64" and st == 5), (b, st)
        assert torch.equal(got[b][0], ref), b


def test_gpu_blocks_carried_through_rebuilds_and_warm_reuse(gpu_lib):
    """Block A runs in handle 1 beside B; handle 2 is built over {A, C}
    (A carried: fresh = 0 and its counters moved), B leaves; later handle 3
    warm-starts B's region with a new right-hand side.  Every result equals
    the solo solve (the warm one from B's previous pressure)."""
    import torch
    from hydro3d_amd.solver import DevicePool, MultigridSolver, block_bytes
    kw = KW["mixed"]
    A, B, C = SPECS[2], SPECS[0], SPECS[1]
    fa, fb, fc, fb2 = _rhs(A, 1), _rhs(B, 2), _rhs(C, 3), _rhs(B, 4)
    print("case: carry through rebuilds, warm reuse", flush=True)
    with DevicePool(2 << 30, 64) as pool:
        oa, ob, oc = (pool.alloc(block_bytes(s, **kw)) for s in (A, B, C))
        with MultigridSolver([A, B], pool=pool, offsets=[oa, ob], tol=1e-8, **kw) as S1:
            S1.set_rhs(0, fa)
            S1.set_rhs(1, fb)
            S1.set_block_tol(0, 1e-11)  # A outlasts B
            S1.solve_begin([0, 1])
            for _ in range(3):
                S1.solve_step()
            sa = S1.block_state(0)
            assert sa.running and sa.cycles == 3
            pb1 = torch.empty_like(fb)
            while S1.block_state(1).running:
                S1.solve_step()
            S1.get_p(1, out=pb1)
            sa = S1.block_state(0)
            assert sa.running  # carried into the next handle mid-solve
        with MultigridSolver([C, A], pool=pool, offsets=[oc, oa], fresh=[1, 0], tol=1e-8,
                             **kw) as S2:
            S2.set_block_state(1, sa)
            S2.set_rhs(0, fc)
            S2.solve_begin([0])
            while S2.solve_step():
                pass
            pa, pc = torch.empty_like(fa), torch.empty_like(fc)
            S2.get_p(1, out=pa)
            S2.get_p(0, out=pc)
            cyc_a, cyc_c = S2.block_state(1).cycles, S2.block_state(0).cycles
        # B's region still holds its pressure: a warm start with a new rhs
        with MultigridSolver([B], pool=pool, offsets=[ob], fresh=[0], tol=1e-8, **kw) as S3:
            S3.set_rhs(0, fb2)
            S3.solve_begin([0])
            while S3.solve_step():
                pass
            pb2 = torch.empty_like(fb2)
            S3.get_p(0, out=pb2)
            cyc_b2 = S3.block_state(0).cycles
        for o in (oa, ob, oc):
            pool.free(o)
        assert pool.free_bytes == pool.nbytes
    ra, ca, _ = _alone(A, fa, kw, 1e-11)
    rb, _, _ = _alone(B, fb, kw, 1e-8)
    rc, cc, _ = _alone(C, fc, kw, 1e-8)
    rb2, cb2, _ = _alone(B, fb2, kw, 1e-8, p0=rb)
    assert torch.equal(pb1, rb)
    assert torch.equal(pa, ra) and cyc_a == ca
    assert torch.equal(pc, rc) and cyc_c == cc
    assert torch.equal(pb2, rb2) and cyc_b2 == cb2


def test_gpu_stepped_outcomes_and_pool_refusals(gpu_lib):
    """maxcy reached -> NOT_CONVERGED in the block's state; a non-finite
    rhs -> DIVERGED for that block only; pooled handles refuse host arrays,
    neighbours and the velocity loop; a pool too small says NOMEM."""
    import torch
    from hydro3d_amd import _lib as L
    from hydro3d_amd.solver import DevicePool, MultigridSolver, block_bytes
    kw = dict(precision="mixed", coarse="exact", maxcy=3, use_graph=False)
    a, b = SPECS[1], SPECS[0]
    print("case: stepped outcomes and refusals", flush=True)
    with DevicePool(1 << 30, 16) as pool:
        offs = [pool.alloc(block_bytes(s, **kw)) for s in (a, b)]
        with MultigridSolver([a, b], pool=pool, offsets=offs, tol=1e-14, **kw) as S:
            S.set_rhs(0, _rhs(a, 7))
            bad = _rhs(b, 8)
            bad[1, 2, 3] = float("nan")
            S.set_rhs(1, bad)
            S.solve_begin([0, 1])
            while S.solve_step():
                pass
            s0, s1 = S.block_state(0), S.block_state(1)
            assert (s0.status, s0.cycles) == (L.ERR_NOT_CONVERGED, 3)
            assert s1.status == L.ERR_DIVERGED
            with pytest.raises(L.H3dError) as e:
                S.get_p(0)  # host array: pooled handles have no staging
            assert e.value.status == L.ERR_UNSUPPORTED
            with pytest.raises(L.H3dError) as e:
                S.solve_begin([0, 0])
            assert e.value.status == L.ERR_ARG
        nb = dict(a, bc=(2, 2, 0, 0, 0, 0), nbr=(0, 0, -1, -1, -1, -1))
        with pytest.raises(L.H3dError) as e:
            MultigridSolver([nb], pool=pool, offsets=[offs[0]], **kw)
        assert e.value.status in (L.ERR_UNSUPPORTED, L.ERR_ARG)
        # a region smaller than the block needs
        small = pool.alloc(1024)
        with pytest.raises(L.H3dError) as e:
            MultigridSolver([b], pool=pool, offsets=[small], **kw)
        assert e.value.status == L.ERR_ARG
        for o in offs + [small]:
            pool.free(o)
        assert pool.free_bytes == pool.nbytes
    with DevicePool(block_bytes(a, **kw), 4) as pool:
        off = pool.alloc(block_bytes(a, **kw))
        with pytest.raises(MemoryError):  # no room left for the handle's tables
            MultigridSolver([a], pool=pool, offsets=[off], **kw)
        pool.free(off)
        assert pool.free_bytes == pool.nbytes
    # a pool with a live handle is not freed under it
    pool = DevicePool(1 << 28, 4)
    off = pool.alloc(block_bytes(a, **kw))
    S = MultigridSolver([a], pool=pool, offsets=[off], **kw)
    pool.lib.h3dmg_pool_destroy(pool.p)
    assert b"live handles" in pool.lib.h3dmg_last_error()
    S.set_rhs(0, _rhs(a, 9))  # its memory is still there
    S.close()
    pool.free(off)
    pool.close()
    torch.cuda.synchronize()
