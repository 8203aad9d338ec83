"""Serving layer: scheduling on the CPU with a host model of the device
backend (admission by footprint, joins into a running batch at the cycle
boundary, residency and warm reuse, LRU eviction, failures); on the GPU,
every served result equals the same request solved alone."""
import math
import threading
import time

import numpy as np
import pytest

from hydro3d_amd import serving as SV


class FakeState:
    def __init__(self, **kw):
        self.running, self.cycles, self.status, self.enabled = 0, 0, -1, 1
        self.fnorm, self.rel, self.tol = 0.0, 0.0, 0.0
        self.__dict__.update(kw)


class FakeHandle:
    """Host model of a pooled handle: a block needs -log10(tol) cycles;
    its pressure is f/2, plus 1 when its region held a pressure already
    (warm); each step takes `step_s`."""

    def __init__(self, backend, specs, offsets, fresh, stream):
        self.be = backend
        self.regions = [backend.mem[o] for o in offsets]
        for r, fr in zip(self.regions, fresh):
            if fr:
                r.clear()
        self.state = [FakeState() for _ in specs]
        self.tol = [1e-8] * len(specs)
        self.closed = False

    def set_block_state(self, b, st):
        self.state[b] = FakeState(**vars(st))

    def block_state(self, b):
        return FakeState(**vars(self.state[b]))

    def set_rhs(self, b, f):
        self.regions[b]["f"] = np.asarray(f)

    def set_block_tol(self, b, tol):
        self.tol[b] = tol

    def solve_begin(self, blocks):
        for b in blocks:
            st = self.state[b]
            st.cycles, st.tol = 0, self.tol[b]
            if not np.all(np.isfinite(self.regions[b]["f"])):
                st.running, st.status = 0, 6
            else:
                st.running, st.status = 1, -1

    def solve_step(self):
        time.sleep(self.be.step_s)
        n = 0
        for b, st in enumerate(self.state):
            if not st.running:
                continue
            st.cycles += 1
            if st.cycles >= round(-math.log10(st.tol)):
                r = self.regions[b]
                r["p"] = r["f"] / 2 + (1.0 if "p" in r else 0.0)
                st.running, st.status, st.rel = 0, 0, st.tol / 2
            n += st.running
        return n

    def result_p(self, b):
        r = self.regions[b]
        return r["p"].copy() if "p" in r else np.zeros_like(r["f"])

    def close(self):
        self.closed = True


class FakeBackend:
    def __init__(self, capacity=1 << 40, step_s=0.002, fail_if=None):
        self.capacity = capacity
        self.step_s = step_s
        self.mem = {}
        self.next = 0
        self.allocs = 0
        self.handles = []
        self.fail_if = fail_if
        self.lock = threading.Lock()

    def block_bytes(self, spec):
        return int(np.prod(spec["n"])) * 8

    def alloc(self, nbytes):
        with self.lock:
            off = self.next
            self.next += nbytes
            self.mem[off] = {}
            self.allocs += 1
            return off

    def free(self, off):
        with self.lock:
            del self.mem[off]

    def new_stream(self):
        return None, 0

    def handle(self, specs, offsets, fresh, stream):
        if self.fail_if and any(self.fail_if(s) for s in specs):
            raise RuntimeError("device refused")
        h = FakeHandle(self, specs, offsets, fresh, stream)
        with self.lock:
            self.handles.append(h)
        return h

    def close(self):
        pass


def _req(n, key=None, tol=1e-8, val=1.0, rid=0):
    return SV.Request(n=n, h=tuple(1.0 / x for x in n), bc=(0,) * 6,
                      f=np.full(n[::-1], val), tol=tol, key=key, rid=rid)


def _server(backend, **kw):
    kw.setdefault("lanes", 1)
    return SV.PoissonServer(backend=backend, **kw)


def test_requests_arriving_during_a_solve_join_the_running_batch():
    be = FakeBackend(step_s=0.005)
    srv = _server(be)
    try:
        f0 = srv.submit(_req((32, 32, 32), key=0, rid=0, tol=1e-8))  # 8 cycles, ~40 ms
        time.sleep(0.012)
        f1 = srv.submit(_req((32, 32, 16), key=1, rid=1, tol=1e-6))
        f2 = srv.submit(_req((16, 32, 32), key=2, rid=2, tol=1e-6))
        r = [f.result(timeout=10) for f in (f0, f1, f2)]
    finally:
        srv.close()
    # the two late requests joined the first one's batch while it ran
    assert srv.stats["joined_running"] >= 1
    assert r[1].batch_blocks > 1.5 and r[0].batch_blocks > 1.0
    assert [x.cycles for x in r] == [8, 6, 6] and all(x.status == 0 for x in r)
    assert all(np.all(x.p == 0.5) for x in r)
    # the running block was carried into the rebuilt handle, not restarted
    assert r[0].latency_s < 0.2


def test_reuse_is_warm_on_the_resident_region():
    be = FakeBackend()
    srv = _server(be)
    try:
        r1 = srv.submit(_req((32, 16, 16), key="sim-a", val=2.0)).result(timeout=10)
        r2 = srv.submit(_req((32, 16, 16), key="sim-a", val=4.0)).result(timeout=10)
    finally:
        srv.close()
    assert not r1.warm and r2.warm
    assert be.allocs == 1  # the second request used the first one's region
    assert np.all(r1.p == 1.0) and np.all(r2.p == 4.0 / 2 + 1)


def test_same_key_requests_queue_behind_each_other():
    be = FakeBackend(step_s=0.003)
    srv = _server(be)
    try:
        fs = [srv.submit(_req((16, 16, 16), key="k", val=2.0 * (q + 1), rid=q)) for q in range(3)]
        rs = [f.result(timeout=10) for f in fs]
    finally:
        srv.close()
    assert [r.warm for r in rs] == [False, True, True]
    assert be.allocs == 1
    assert np.all(rs[2].p == 3.0 + 1)


def test_budget_evicts_idle_regions_lru():
    one = 64 ** 3 * 8
    be = FakeBackend(capacity=int(2.5 * one))
    srv = _server(be)
    try:
        for i in range(4):
            srv.submit(_req((64, 64, 64), key=i, rid=i)).result(timeout=10)
        assert srv.resident_bytes <= int(2.5 * one)
        assert srv.stats["evicted"] >= 2
        assert 0 not in srv._by_key and 3 in srv._by_key
    finally:
        srv.close()


def test_request_waits_for_memory_held_by_a_running_solve():
    one = 32 ** 3 * 8
    be = FakeBackend(capacity=int(1.5 * one), step_s=0.01)
    srv = _server(be, lanes=2, lane_cells=[32 ** 3 - 1])
    try:
        fa = srv.submit(_req((32, 32, 32), key="a", rid=0))  # 8 cycles ~ 80 ms
        time.sleep(0.005)
        fb = srv.submit(_req((32, 32, 32), key="b", rid=1))
        ra, rb = fa.result(timeout=10), fb.result(timeout=10)
    finally:
        srv.close()
    assert ra.status == rb.status == 0
    assert rb.latency_s > 0.07  # admitted once the first solve gave its region back
    assert srv.stats["evicted"] >= 1


def test_request_larger_than_budget_fails_cleanly():
    be = FakeBackend(capacity=1 << 20)
    srv = _server(be)
    try:
        with pytest.raises(MemoryError):
            srv.submit(_req((64, 64, 64))).result(timeout=10)
        ok = srv.submit(_req((16, 16, 16))).result(timeout=10)
        assert ok.status == 0
    finally:
        srv.close()


def test_handle_build_failure_fails_only_the_joiners():
    be = FakeBackend(step_s=0.005, fail_if=lambda s: s["n"][0] == 48)
    srv = _server(be)
    try:
        good = srv.submit(_req((32, 16, 16), key="good", rid=1))
        time.sleep(0.01)
        bad = srv.submit(_req((48, 16, 16), key="bad", rid=2))
        with pytest.raises(RuntimeError, match="device refused"):
            bad.result(timeout=10)
        assert good.result(timeout=10).status == 0
        later = srv.submit(_req((16, 16, 16), key="x", rid=3)).result(timeout=10)
        assert later.status == 0
        assert "bad" not in srv._by_key
        assert srv.resident_bytes == 16 ** 3 * 8 + 32 * 16 * 16 * 8
    finally:
        srv.close()
    assert srv.resident_bytes == 0 and not be.mem


def test_nonfinite_rhs_is_diverged_at_once_and_alone():
    be = FakeBackend(step_s=0.003)
    srv = _server(be)
    try:
        f = np.ones((16, 16, 16))
        f[3, 4, 5] = np.nan
        ok = srv.submit(_req((16, 16, 32), rid=0))
        bad = srv.submit(SV.Request(n=(16, 16, 16), h=(1,) * 3, bc=(0,) * 6, f=f, rid=1))
        rb, ro = bad.result(timeout=10), ok.result(timeout=10)
    finally:
        srv.close()
    assert rb.status == 6 and rb.cycles == 0
    assert ro.status == 0 and ro.cycles == 8


def test_lanes_split_by_size_class():
    srv = SV.PoissonServer(backend=FakeBackend(), lanes=3)
    try:
        small = dict(n=(32, 32, 32))
        mid = dict(n=(96, 96, 96))
        big = dict(n=(256, 256, 128))
        assert [srv._lane_of(x) for x in (small, mid, big)] == [0, 1, 2]
    finally:
        srv.close()


def test_batch_is_capped():
    be = FakeBackend(step_s=0.003)
    srv = _server(be, max_batch_blocks=2)
    try:
        fs = [srv.submit(_req((16, 16, 16), key=q, rid=q, tol=1e-6)) for q in range(5)]
        rs = [f.result(timeout=10) for f in fs]
    finally:
        srv.close()
    assert all(r.status == 0 for r in rs)
    assert max(r.batch_blocks for r in rs) <= 2.0


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
    """Requests of different sizes (up to 256 per axis), boundaries and
    tolerances served by the continuous-batching server -- joining running
    batches, warm reuse of resident regions -- give bitwise the pressure each
    gets from a fresh solver alone."""
    import torch
    from hydro3d_amd import MultigridSolver
    torch.cuda.set_device(0)
    # arrivals faster than the configured 200/s, so that requests join
    # batches that are running
    reqs = SV.request_stream(160, rate=1000.0, seed=11, hi=256, reuse_frac=0.3)
    print("case: 160 served requests (sizes 32..256) vs alone", flush=True)
    with SV.PoissonServer(hbm_budget=16 << 30, lanes=3) as srv:
        results, _ = SV.run_stream(srv, reqs)
        stats = dict(srv.stats)
    print("server", stats, flush=True)
    assert stats["joined_running"] > 0
    assert max(r.batch_blocks for r in results) > 1.5
    assert any(r.warm for r in results)
    by_key = {}
    for d, r in zip(reqs, results):
        assert r.status == 0, (d["rid"], r.status)
        f = SV.device_rhs(d)
        spec = SV.Request(n=d["n"], h=d["h"], bc=d["bc"], f=None).spec()
        p0 = by_key[d["key"]] if r.warm else torch.zeros_like(f)
        with MultigridSolver([spec], tol=d["tol"], precision="mixed", coarse="exact") as S:
            S.set_rhs(0, f)
            S.set_p(0, p0)
            S.solve()
            alone = S.result_p(0)
            cyc, _ = S.block_stats()
        assert torch.equal(alone, r.p), d["rid"]
        assert r.cycles == cyc[0], d["rid"]
        by_key[d["key"]] = r.p


@pytest.mark.gpu
def test_pooled_tables_fit_the_lane_reserve():
    """The pool holds back DeviceBackend.RESERVE_PER_LANE per lane for the
    handles' own tables: a full batch of the largest requests fits in it."""
    from hydro3d_amd.solver import DevicePool, MultigridSolver, block_bytes
    print("case: pooled tables vs reserve", flush=True)
    kw = dict(precision="mixed", coarse="exact", use_graph=False)
    big = dict(n=(256, 256, 256), h=(1 / 64,) * 3, bc=(1, 1, 0, 0, 1, 1), nlev=6)
    small = dict(n=(32, 48, 32), h=(1 / 64,) * 3, bc=(0,) * 6, nlev=4)
    specs = [big] * 4 + [small] * 28
    nb = block_bytes(big, **kw), block_bytes(small, **kw)
    with DevicePool(4 * nb[0] + 28 * nb[1] + SV.DeviceBackend.RESERVE_PER_LANE, 64) as pool:
        offs = [pool.alloc(nb[0] if s is big else nb[1]) for s in specs]
        before = pool.free_bytes
        with MultigridSolver(specs, pool=pool, offsets=offs, **kw):
            tables = before - pool.free_bytes
        assert pool.free_bytes == before
        assert 0 < tables < SV.DeviceBackend.RESERVE_PER_LANE // 2, tables
        for o in offs:
            pool.free(o)
