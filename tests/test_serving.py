"""Serving layer: scheduling on the CPU with a host model of the solver
(admission by footprint, batching, residency and warm reuse, LRU eviction);
on the GPU, every served result equals the same request solved alone."""
import threading
import time

import numpy as np
import pytest

from hydro3d_amd import serving as SV


class FakeSolver:
    """Host stand-in with the solver's control surface; 'solves' take a
    little time and return f/2 as the pressure."""
    created = []
    lock = threading.Lock()

    def __init__(self, specs, tol=1e-8, **kw):
        self.blocks = [dict(s) for s in specs]
        self.enabled = [True] * len(specs)
        self.f = [None] * len(specs)
        self.tol = [tol] * len(specs)
        self.p = [None] * len(specs)
        self.closed = False
        with FakeSolver.lock:
            FakeSolver.created.append(self)

    def set_block_enabled(self, b, on):
        self.enabled[b] = bool(on)

    def set_rhs(self, b, f):
        self.f[b] = np.asarray(f)

    def set_block_tol(self, b, tol):
        self.tol[b] = tol

    def solve(self):
        time.sleep(0.01)
        for b, on in enumerate(self.enabled):
            if on:
                warm = self.p[b] is not None
                self.p[b] = self.f[b] / 2 + (1 if warm else 0)
        return 0, 3, None

    def block_stats(self):
        return [3] * len(self.blocks), [t / 2 for t in self.tol]

    def result_p(self, b):
        return self.p[b]

    def close(self):
        self.closed = True


def _req(n, key=None, tol=1e-8, val=1.0, rid=0):
    return SV.Request(n=n, h=tuple(1.0 / x for x in n), bc=(0,) * 6,
                      f=np.full(n[::-1], val), tol=tol, key=key, rid=rid)


@pytest.fixture(autouse=True)
def _reset():
    FakeSolver.created = []


def test_footprint_grows_with_blocks_and_precision():
    a = dict(n=(64, 64, 64), h=(1,) * 3, bc=(0,) * 6, nlev=4)
    b = dict(n=(128, 64, 32), h=(1,) * 3, bc=(0,) * 6, nlev=4)
    f1 = SV.footprint_bytes([a])
    f2 = SV.footprint_bytes([b])
    # the staging array is shared by the blocks of a handle
    assert max(f1, f2) < SV.footprint_bytes([a, b]) <= f1 + f2
    assert SV.footprint_bytes([a], "fp64", "smooth") > f1
    # 256^3 mixed: the arena is ~0.64 GB
    big = dict(n=(256,) * 3, h=(1,) * 3, bc=(0,) * 6, nlev=6)
    assert 0.6e9 < SV.footprint_bytes([big]) < 0.8e9


def test_requests_waiting_together_share_one_batch():
    srv = SV.PoissonServer(max_batches=1, solver_factory=FakeSolver)
    try:
        futs = [srv.submit(_req((32, 32, 32), key=i, rid=i)) for i in range(6)]
        res = [f.result(timeout=10) for f in futs]
    finally:
        srv.close()
    # the first batch starts at once; the other five wait and go together
    assert sum(r.batch_blocks for r in res) >= 6
    assert max(r.batch_blocks for r in res) >= 2
    assert len(FakeSolver.created) < 6
    assert all(np.all(r.p == 0.5) for r in res)


def test_reuse_is_warm_and_on_the_resident_handle():
    srv = SV.PoissonServer(solver_factory=FakeSolver)
    try:
        r1 = srv.submit(_req((32, 16, 16), key="sim-a", val=2.0)).result(timeout=10)
        r2 = srv.submit(_req((32, 16, 16), key="sim-a", val=4.0)).result(timeout=10)
    finally:
        srv.close()
    assert not r1.warm and r2.warm
    assert len(FakeSolver.created) == 1
    assert np.all(r2.p == 4.0 / 2 + 1)  # warm: the handle kept its pressure


def test_budget_evicts_idle_hierarchies_lru():
    one = SV.footprint_bytes([dict(n=(64,) * 3, h=(1,) * 3, bc=(0,) * 6, nlev=4)])
    srv = SV.PoissonServer(hbm_budget=int(2.5 * one), max_batches=1, solver_factory=FakeSolver)
    try:
        for i in range(4):
            srv.submit(_req((64, 64, 64), key=i, rid=i)).result(timeout=10)
        assert srv.resident_bytes <= int(2.5 * one)
        assert srv.stats["evicted"] >= 2
        assert FakeSolver.created[0].closed and not FakeSolver.created[-1].closed
    finally:
        srv.close()


def test_request_larger_than_budget_fails_cleanly():
    srv = SV.PoissonServer(hbm_budget=1 << 20, solver_factory=FakeSolver)
    try:
        fut = srv.submit(_req((64, 64, 64)))
        with pytest.raises(MemoryError):
            fut.result(timeout=10)
    finally:
        srv.close()


def test_handle_build_failure_fails_its_batch_and_queued_reuse():
    """Handles are built by the worker thread (the footprint is reserved
    first).  A build that raises fails its batch and any request queued for
    its key meanwhile; the server keeps serving."""
    gate = threading.Event()

    class SlowFailing(FakeSolver):
        def __init__(self, specs, **kw):
            if specs[0]["n"][0] == 48:
                gate.wait(5)  # let the same key's second request queue up
                raise RuntimeError("device refused")
            super().__init__(specs, **kw)

    srv = SV.PoissonServer(solver_factory=SlowFailing, max_batch_blocks=1)
    try:
        f1 = srv.submit(_req((48, 16, 16), key="bad", rid=1))
        time.sleep(0.05)
        f2 = srv.submit(_req((48, 16, 16), key="bad", rid=2))
        gate.set()
        for f in (f1, f2):
            with pytest.raises(RuntimeError, match="device refused"):
                f.result(timeout=10)
        ok = srv.submit(_req((32, 16, 16), key="good", rid=3)).result(timeout=10)
        assert ok.status == 0 and srv.resident_bytes > 0
    finally:
        srv.close()
    assert "bad" not in srv._by_key


def test_handles_are_built_concurrently():
    """Two new hierarchies whose builds overlap in time: the dispatcher does
    not hold the lock across a build."""
    inside, peak = [0], [0]
    lk = threading.Lock()

    class Slow(FakeSolver):
        def __init__(self, specs, **kw):
            with lk:
                inside[0] += 1
                peak[0] = max(peak[0], inside[0])
            time.sleep(0.2)
            with lk:
                inside[0] -= 1
            super().__init__(specs, **kw)

    srv = SV.PoissonServer(solver_factory=Slow, max_batch_blocks=1, max_batches=4)
    try:
        futs = [srv.submit(_req((32 + 16 * i, 16, 16), rid=i)) for i in range(3)]
        for f in futs:
            f.result(timeout=10)
    finally:
        srv.close()
    assert peak[0] >= 2


def test_block_status_judges_each_block_by_its_own_residual():
    """A multi-block solve returns NOT_CONVERGED (5) when any block reaches
    maxcy; blocks that met their own tolerance are still converged."""
    assert SV.block_status(5, 3e-9, 1e-8) == 0
    assert SV.block_status(5, 2e-8, 1e-8) == 5
    assert SV.block_status(0, 1e-9, 1e-8) == 0
    assert SV.block_status(5, float("nan"), 1e-8) == 5
    # a diverged or failed solve speaks for every block
    assert SV.block_status(6, 1e-12, 1e-8) == 6
    assert SV.block_status(3, 1e-12, 1e-8) == 3


def test_served_status_is_per_block():
    class OneSlow(FakeSolver):
        def solve(self):
            super().solve()
            return 5, 60, None  # the batch as a whole hit maxcy

        def block_stats(self):
            # block 0 met its tolerance, every other block did not
            n = len(self.blocks)
            return [7] + [60] * (n - 1), [self.tol[0] / 2] + [t * 3 for t in self.tol[1:]]

    gate = threading.Event()

    class Gated(OneSlow):
        def __init__(self, specs, **kw):
            gate.wait(5)
            super().__init__(specs, **kw)

    srv = SV.PoissonServer(max_batches=1, solver_factory=Gated)
    try:
        f0 = srv.submit(_req((32, 32, 32), rid=0))
        f1 = srv.submit(_req((32, 32, 32), rid=1))
        gate.set()
        r0, r1 = f0.result(timeout=10), f1.result(timeout=10)
    finally:
        srv.close()
    # whichever batch shape the dispatcher chose, each result is judged alone
    for r in (r0, r1):
        assert r.status in (0, 5)
    if r0.batch_blocks == 2:
        assert (r0.status, r1.status) == (0, 5)


def test_request_stream_follows_the_recipe():
    reqs = SV.request_stream(400, rate=200.0, seed=4)
    t = np.array([r["t"] for r in reqs])
    assert np.all(np.diff(t) > 0) and 1.5 < t[-1] < 2.5  # ~400 / 200 s
    sizes = np.array([r["n"] for r in reqs])
    assert sizes.min() >= 32 and sizes.max() <= 256 and np.all(sizes % 16 == 0)
    reuse = sum(1 for r in reqs if r["step"] > 0)
    assert 0.12 < reuse / len(reqs) < 0.28
    assert {r["tol"] for r in reqs} == {1e-6, 1e-8}


@pytest.mark.gpu
def test_served_results_equal_solving_alone():
    """Requests of different sizes, boundaries and tolerances served
    concurrently (batched, overlapping, one warm reuse) give bitwise the
    pressure each gets from a fresh solver alone."""
    import torch
    from hydro3d_amd import MultigridSolver
    torch.cuda.set_device(0)
    reqs = SV.request_stream(120, rate=200.0, seed=11, hi=256)
    print("case: 120 served requests (sizes 32..256) vs alone", flush=True)
    with SV.PoissonServer(max_batches=3) as srv:
        results, _ = SV.run_stream(srv, reqs)
    by_key = {}
    for d, r in zip(reqs, results):
        assert r.status == 0, (d["rid"], r.status)
        f = SV.device_rhs(d)
        spec = SV.Request(n=d["n"], h=d["h"], bc=d["bc"], f=None).spec()
        if r.warm:
            # alone: the same warm start (the key's previous pressure)
            p0 = by_key[d["key"]]
        else:
            p0 = torch.zeros_like(f)
        with MultigridSolver([spec], tol=d["tol"], precision="mixed", coarse="exact") as S:
            S.set_rhs(0, f)
            S.set_p(0, p0)
            S.solve()
            alone = S.result_p(0)
        assert torch.equal(alone, r.p), d["rid"]
        by_key[d["key"]] = r.p


@pytest.mark.gpu
def test_footprint_covers_the_device_arena():
    """Admission uses footprint_bytes: it must not undercount what a handle
    allocates."""
    from hydro3d_amd import MultigridSolver
    print("case: footprint vs device bytes", flush=True)
    shapes = [[dict(n=(64, 64, 64), h=(1 / 64,) * 3, bc=(1, 1, 1, 1, 0, 0), nlev=4)],
              [dict(n=(96, 48, 32), h=(0.01,) * 3, bc=(0,) * 6, nlev=4),
               dict(n=(32, 32, 32), h=(0.03,) * 3, bc=(1,) * 6, nlev=3)]]
    for specs in shapes:
        for prec, coarse in (("mixed", "exact"), ("fp64", "smooth")):
            with MultigridSolver(specs, precision=prec, coarse=coarse) as S:
                dev = S.device_bytes
            est = SV.footprint_bytes(specs, prec, coarse)
            assert dev <= est <= 1.5 * dev + (4 << 20), (specs, prec, dev, est)
