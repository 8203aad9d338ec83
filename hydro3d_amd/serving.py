"""Serving: continuous batching of Poisson solves from concurrent simulations.

Requests (one block each: size, spacing, boundaries, right-hand side,
tolerance, optional residency key) arrive at any time.  The server

* admits a request by HBM footprint: the block's region (its hierarchy,
  pressure and coarse-solve tables, h3dmg_block_bytes) is cut from one
  device pool made at start-up.  When the pool is full, idle resident
  regions are evicted least-recently-used first; a request that still does
  not fit waits until running solves leave, and one larger than the whole
  pool fails at once;
* runs `lanes` batches at once, each a pooled handle on its own host thread
  and HIP stream; a request goes to the lane of its size class (cells), so
  small solves do not step in lock-step with the largest ones;
* admits waiting requests INTO a running batch at the next V-cycle
  boundary: the lane builds a new handle over the running blocks plus the
  newcomers (the running blocks keep their state in their regions,
  h3dmg_create_pooled with fresh = 0, their counters carried), starts the
  newcomers (h3dmg_solve_begin) and steps every running block one V-cycle
  per launch sequence (h3dmg_solve_step); each block leaves when it meets
  its own tolerance;
* keeps a finished block's region -- hierarchy and pressure -- resident
  under the request's key: a later request with the same key joins with
  that region, warm-started from the previous pressure.

Blocks of one handle never interact (per-block norms, tolerances and
stopping), so each result is bitwise the one the request gets alone.
Nothing is allocated on the request path: device memory comes from the pool
(block regions and every handle's tables), the per-block statistics from the
pool's pinned array; only the result tensor comes from torch's caching
allocator.

    python -m hydro3d_amd.serving --requests 2000 --rate 200 --seed 4
"""
from __future__ import annotations

import argparse
import collections
import json
import math
import sys
import threading
import time
from concurrent.futures import Future
from dataclasses import dataclass, field

import numpy as np

from .configs import levels_for


# ---- requests and results -------------------------------------------------
@dataclass
class Request:
    n: tuple
    h: tuple
    bc: tuple
    f: object                 # device tensor (nk, nj, ni) float64, or numpy
    tol: float = 1e-8
    key: object = None        # residency key: same key = same simulation
    nlev: int = 0             # 0: min(6, v2(n))
    rid: int = -1
    t_submit: float = 0.0

    def spec(self):
        return dict(n=tuple(self.n), h=tuple(self.h), bc=tuple(self.bc),
                    nlev=self.nlev or levels_for(self.n))


@dataclass
class Result:
    rid: int
    p: object                 # device tensor (nk, nj, ni): the pressure
    status: int               # 0 converged, 5 maxcy reached, 6 diverged
    cycles: int
    rel_residual: float
    latency_s: float
    batch_blocks: float       # blocks stepping beside it, mean over its cycles (itself included)
    warm: bool


@dataclass
class _Region:
    """A block's piece of the pool: its hierarchy and pressure."""
    off: int
    nbytes: int
    spec: dict
    key: object
    busy: bool = True         # queued for or held by a lane
    fresh: bool = True        # no handle has used it yet: starts from zero
    last_used: float = field(default_factory=time.perf_counter)


@dataclass
class _Job:
    req: Request
    fut: Future
    region: _Region
    warm: bool
    b: int = -1               # index in the lane's current handle
    running: bool = True
    batch_sum: int = 0
    steps: int = 0


def _same_block(a, b):
    return (tuple(a["n"]) == tuple(b["n"]) and tuple(a["bc"]) == tuple(b["bc"])
            and tuple(a["h"]) == tuple(b["h"]) and a["nlev"] == b["nlev"])


# ---- device backend -------------------------------------------------------
class DeviceBackend:
    """The pool and pooled handles of the device library.  The server only
    needs this surface; tests substitute a host model of it."""

    # device memory held back for the handles' own tables (per lane)
    RESERVE_PER_LANE = 256 << 20

    def __init__(self, nbytes, lanes, max_blocks, params):
        from .solver import DevicePool
        self.params = params
        self.pool = DevicePool(nbytes, max_blocks)
        self.capacity = nbytes - lanes * self.RESERVE_PER_LANE
        self._bytes = {}

    def block_bytes(self, spec):
        from .solver import block_bytes
        k = (tuple(spec["n"]), tuple(spec["bc"]), tuple(spec["h"]), spec["nlev"])
        if k not in self._bytes:
            self._bytes[k] = block_bytes(spec, precision=self.params["precision"],
                                         coarse=self.params["coarse"])
        return self._bytes[k]

    def alloc(self, nbytes):
        return self.pool.alloc(nbytes)  # MemoryError when no piece is free

    def free(self, off):
        self.pool.free(off)

    def new_stream(self):
        import torch
        s = torch.cuda.Stream()
        return s, s.cuda_stream

    def handle(self, specs, offsets, fresh, stream):
        from .solver import MultigridSolver
        return MultigridSolver(specs, pool=self.pool, offsets=offsets, fresh=fresh,
                               stream=stream, **self.params)

    def close(self):
        self.pool.close()


# ---- the server -----------------------------------------------------------
def default_lane_cells(lanes):
    """Upper cell counts of the lanes' size classes: 64^3 and 128^3 for
    three lanes (a quarter, half and a quarter of the stream's requests)."""
    if lanes <= 1:
        return []
    if lanes == 2:
        return [96 ** 3]
    edges = np.geomspace(64 ** 3, 128 ** 3, lanes - 1)
    return [int(x) for x in edges]


class PoissonServer:
    """Continuous-batching server over pooled device handles."""

    def __init__(self, hbm_budget=32 << 30, lanes=3, max_batch_blocks=32, precision="mixed",
                 coarse="exact", nu1=2, nu2=2, maxcy=60, keep_resident=True, use_graph=False,
                 lane_cells=None, backend=None):
        self.params = dict(nu1=nu1, nu2=nu2, maxcy=maxcy, precision=precision, coarse=coarse,
                           use_graph=use_graph)
        self.lanes = int(lanes)
        self.max_batch_blocks = int(max_batch_blocks)
        self.keep_resident = keep_resident
        self.lane_cells = list(default_lane_cells(self.lanes) if lane_cells is None else lane_cells)
        assert len(self.lane_cells) == self.lanes - 1
        self.backend = backend or DeviceBackend(int(hbm_budget), self.lanes,
                                                4 * self.lanes * self.max_batch_blocks + 64,
                                                self.params)
        self.capacity = self.backend.capacity
        self._lock = threading.Condition()
        self._admit = collections.deque()                    # (Request, Future) waiting
        self._queues = [collections.deque() for _ in range(self.lanes)]  # admitted _Jobs
        self._by_key = {}                                    # key -> _Region
        self._idle = collections.OrderedDict()               # id(_Region) -> _Region, LRU first
        self._used = 0                                       # bytes of live regions
        self._closed = False
        self.stats = collections.Counter()
        self._threads = [threading.Thread(target=self._lane, args=(q,), daemon=True,
                                          name=f"h3dmg-lane{q}") for q in range(self.lanes)]
        for t in self._threads:
            t.start()

    # -- public
    def submit(self, req: Request) -> Future:
        fut = Future()
        req.t_submit = req.t_submit or time.perf_counter()
        with self._lock:
            if self._closed:
                raise RuntimeError("server closed")
            self._admit.append((req, fut))
            self._try_admit()
            self._lock.notify_all()
        return fut

    def warmup(self, n=(32, 32, 32)):
        """Serve one small request and forget it: the kernels' code objects
        and the device context are loaded before traffic arrives.  The
        counters start from zero afterwards."""
        d = dict(n=tuple(n), h=(1.0 / 64,) * 3, bc=(0,) * 6, tol=1e-6, rid=-1, t=0.0,
                 seed=0)
        req = Request(n=d["n"], h=d["h"], bc=d["bc"], f=device_rhs(d), tol=d["tol"], rid=-1)
        self.submit(req).result()
        with self._lock:
            self.stats.clear()

    def close(self):
        with self._lock:
            self._closed = True
            self._lock.notify_all()
        for t in self._threads:
            t.join()
        with self._lock:
            while self._admit:  # could never be admitted
                _, fut = self._admit.popleft()
                fut.set_exception(MemoryError("server closed before the request fit"))
            for qu in self._queues:
                while qu:
                    job = qu.popleft()
                    job.fut.set_exception(RuntimeError("server closed"))
                    self._release(job)
            for r in list(self._idle.values()):
                self._free_region(r)
            self._by_key.clear()
        self.backend.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    @property
    def resident_bytes(self):
        with self._lock:
            return self._used

    # -- admission (lock held)
    def _lane_of(self, spec):
        cells = spec["n"][0] * spec["n"][1] * spec["n"][2]
        for q, top in enumerate(self.lane_cells):
            if cells <= top:
                return q
        return self.lanes - 1

    def _free_region(self, r):
        self._idle.pop(id(r), None)
        if r.key is not None and self._by_key.get(r.key) is r:
            del self._by_key[r.key]
        self.backend.free(r.off)
        self._used -= r.nbytes

    def _take_memory(self, nbytes):
        """Pool offset of a new region, evicting idle resident regions (LRU)
        as needed; None when it does not fit now."""
        while True:
            if self._used + nbytes <= self.capacity:
                try:
                    off = self.backend.alloc(nbytes)
                    self._used += nbytes
                    return off
                except MemoryError:
                    pass  # fragmented: evict and retry
            if not self._idle:
                return None
            _, victim = self._idle.popitem(last=False)
            self._free_region(victim)
            self.stats["evicted"] += 1

    def _try_admit(self):
        blocked = False
        keep = collections.deque()
        while self._admit:
            req, fut = self._admit.popleft()
            if blocked:
                keep.append((req, fut))
                continue
            spec = req.spec()
            hit = self._by_key.get(req.key) if req.key is not None else None
            if hit is not None and _same_block(hit.spec, spec):
                if hit.busy:  # the key's previous request still holds it
                    keep.append((req, fut))
                    continue
                hit.busy = True
                self._idle.pop(id(hit), None)
                self._queues[self._lane_of(spec)].append(_Job(req, fut, hit, warm=True))
                continue
            try:
                nbytes = self.backend.block_bytes(spec)
            except Exception as e:  # noqa: BLE001 - invalid block: fail the request
                fut.set_exception(e)
                continue
            if nbytes > self.capacity:
                fut.set_exception(MemoryError(f"request needs {nbytes} bytes, "
                                              f"the pool holds {self.capacity}"))
                continue
            off = self._take_memory(nbytes)
            if off is None:
                blocked = True  # first come, first served: later requests wait too
                keep.append((req, fut))
                continue
            region = _Region(off, nbytes, spec, req.key)
            if req.key is not None:
                old = self._by_key.get(req.key)
                if old is not None and not old.busy:
                    self._free_region(old)
                self._by_key[req.key] = region
            self._queues[self._lane_of(spec)].append(_Job(req, fut, region, warm=False))
        self._admit = keep

    def _release(self, job):
        """A finished job's region leaves its handle (lock held)."""
        r = job.region
        r.busy = False
        r.last_used = time.perf_counter()
        if self.keep_resident and r.key is not None and self._by_key.get(r.key) is r:
            self._idle[id(r)] = r
        else:
            self._free_region(r)

    # -- lanes
    def _lane(self, q):
        """One lane: a pooled handle stepped one V-cycle at a time; waiting
        requests join at the cycle boundary, finished blocks leave at once
        (their regions go back: every stepping kernel skips a block that is
        not running, so the handle never touches a region it no longer
        owns).  A device error fails the lane's requests in flight; the lane
        goes on with the next ones."""
        lane = dict(S=None, running=[], joiners=[], stream=None)
        queue = self._queues[q]
        while True:
            try:
                if not self._lane_iteration(lane, queue):
                    break
            except Exception as e:  # noqa: BLE001
                with self._lock:
                    for m in lane["running"] + lane["joiners"]:
                        if not m.fut.done():
                            m.fut.set_exception(e)
                        if m.region.busy:
                            m.region.busy = False
                            self._free_region(m.region)  # state unknown: not resident
                    self._try_admit()
                    self._lock.notify_all()
                lane["running"], lane["joiners"] = [], []
                if lane["S"] is not None:
                    try:
                        lane["S"].close()
                    except Exception:  # noqa: BLE001
                        pass
                    lane["S"] = None
        if lane["S"] is not None:
            lane["S"].close()

    def _lane_iteration(self, lane, queue):
        """One V-cycle boundary of a lane; False when the lane is done."""
        S, running = lane["S"], lane["running"]
        with self._lock:
            while not self._closed and not queue and not running and S is None:
                self._lock.wait()
            if self._closed and not queue and not running and S is None and not self._admit:
                return False
            if not queue and not running and S is None:
                # closing, but requests still wait for memory or for their
                # key: another lane's finish admits them
                self._lock.wait(0.01)
                return True
            room = self.max_batch_blocks - len(running)
            lane["joiners"] = joiners = [queue.popleft() for _ in range(min(room, len(queue)))]
        if not joiners and not running:
            S.close()  # idle
            lane["S"] = None
            return True
        if lane["stream"] is None:
            lane["stream"] = self.backend.new_stream()
        if joiners:
            S, running = self._rebuild(lane, joiners)
            lane["S"], lane["running"], lane["joiners"] = S, running, []
            if not running:
                return True
        S.solve_step()
        self.stats["steps"] += 1
        self.stats["block_steps"] += len(running)
        now, left = None, []
        for m in running:
            m.batch_sum += len(running)
            m.steps += 1
        for m in running:
            st = S.block_state(m.b)
            if st.running:
                left.append(m)
                continue
            now = now or time.perf_counter()
            self._finish(S, m, st, now)
        lane["running"] = left
        return True

    def _finish(self, S, job, st, now=None):
        p = S.result_p(job.b)
        done = now or time.perf_counter()
        job.running = False
        job.fut.set_result(Result(
            rid=job.req.rid, p=p, status=int(st.status), cycles=int(st.cycles),
            rel_residual=float(st.rel), latency_s=done - job.req.t_submit,
            batch_blocks=(job.batch_sum / job.steps) if job.steps else 1.0, warm=job.warm))
        with self._lock:
            self.stats["solves"] += 1
            self.stats["warm"] += 1 if job.warm else 0
            self._release(job)
            self._try_admit()
            self._lock.notify_all()

    def _rebuild(self, lane, joiners):
        """At the V-cycle boundary: a new handle over the running blocks
        (state carried in their regions and counters) and the joiners."""
        S, running = lane["S"], lane["running"]
        t0 = time.perf_counter()
        states = [S.block_state(m.b) for m in running]
        new = running + joiners
        try:
            S2 = self.backend.handle([m.region.spec for m in new], [m.region.off for m in new],
                                     [False] * len(running) + [j.region.fresh for j in joiners],
                                     lane["stream"][1])
        except Exception as e:  # noqa: BLE001 - no handle: the joiners fail, the batch runs on
            with self._lock:
                for j in joiners:
                    j.running = False
                    j.fut.set_exception(e)
                    j.region.busy = False
                    self._free_region(j.region)  # its state is unknown: not resident
                self._try_admit()
                self._lock.notify_all()
            return S, running
        t1 = time.perf_counter()
        if S is not None:
            S.close()
        lane["S"] = S2  # from here on, an error closes S2
        for i, m in enumerate(new):
            m.b = i
        for i, m in enumerate(running):
            S2.set_block_state(i, states[i])
        t2 = time.perf_counter()
        for j in joiners:
            S2.set_rhs(j.b, j.req.f)
            S2.set_block_tol(j.b, j.req.tol)
            j.region.fresh = False
        t3 = time.perf_counter()
        S2.solve_begin([j.b for j in joiners])
        t4 = time.perf_counter()
        with self._lock:
            # host and device time of the rebuild and the newcomers' start:
            # the new handle, the old one's release, the right-hand sides,
            # the newcomers' first residual
            self.stats["rebuild_us"] += int((t4 - t0) * 1e6)
            self.stats["rebuild_create_us"] += int((t1 - t0) * 1e6)
            self.stats["rebuild_close_us"] += int((t2 - t1) * 1e6)
            self.stats["rebuild_rhs_us"] += int((t3 - t2) * 1e6)
            self.stats["rebuild_begin_us"] += int((t4 - t3) * 1e6)
            self.stats["rebuilds"] += 1
            self.stats["joined"] += len(joiners)
            self.stats["joined_running"] += len(joiners) if running else 0
        left = list(running)
        for j in joiners:
            st = S2.block_state(j.b)
            if st.running:
                left.append(j)
            else:  # finished at the start (a non-finite rhs)
                self._finish(S2, j, st)
        return S2, left


# ---- the serving workload: seeded request stream --------------------------
def request_stream(nreq, rate=200.0, seed=4, reuse_frac=0.2, lo=32, hi=256, quantum=16):
    """Arrival offsets (s) and request descriptors: exponential gaps of mean
    1/rate; per-axis sizes exp(U(ln lo, ln hi)) rounded to a multiple of
    `quantum`, cubic cells of side 1/64 (the spacing of the single-block and
    batched workloads); per-axis Neumann or periodic by seed; tolerance 1e-6
    or 1e-8;
    a fraction `reuse_frac` repeats an earlier request's key (same geometry,
    next right-hand side of that simulation, warm start)."""
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate, size=nreq))
    out = []
    for i in range(nreq):
        if out and rng.random() < reuse_frac:
            src = out[int(rng.integers(len(out)))]
            d = dict(src, rid=i, t=float(t[i]), tol=float(rng.choice([1e-6, 1e-8])),
                     step=src["step"] + 1, seed=int(rng.integers(1 << 31)))
            out.append(d)
            continue
        n = tuple(int(max(quantum, round(math.exp(rng.uniform(math.log(lo), math.log(hi))) / quantum)
                          * quantum)) for _ in range(3))
        per = [bool(rng.random() < 0.5) for _ in range(3)]
        bc = tuple(x for a in range(3) for x in ((1, 1) if per[a] else (0, 0)))
        out.append(dict(rid=i, t=float(t[i]), n=n, bc=bc, h=(1.0 / 64,) * 3,
                        tol=float(rng.choice([1e-6, 1e-8])), key=i, step=0,
                        seed=int(rng.integers(1 << 31))))
    return out


def device_rhs(d, dt=1e-3):
    """Right-hand side of a stream request on the GPU: div(u)/dt of N(0,1)
    face velocities (periodic wrap or no-penetration walls), mean removed --
    the recipe of hydro3d_amd.configs, drawn with torch's generator so that
    large requests are made at arrival rate."""
    import torch
    ni, nj, nk = d["n"]
    g = torch.Generator(device="cuda")
    g.manual_seed(d["seed"])
    u = torch.randn((nk, nj, ni + 1), generator=g, device="cuda", dtype=torch.float64)
    v = torch.randn((nk, nj + 1, ni), generator=g, device="cuda", dtype=torch.float64)
    w = torch.randn((nk + 1, nj, ni), generator=g, device="cuda", dtype=torch.float64)
    for arr, ax, lo in ((u, 2, d["bc"][0]), (v, 1, d["bc"][2]), (w, 0, d["bc"][4])):
        first = [slice(None)] * 3
        last = [slice(None)] * 3
        first[ax], last[ax] = 0, arr.shape[ax] - 1
        if lo == 1:
            arr[tuple(first)] = arr[tuple(last)]
        else:
            arr[tuple(first)] = 0.0
            arr[tuple(last)] = 0.0
    dx, dy, dz = d["h"]
    f = ((u[:, :, 1:] - u[:, :, :-1]) / dx + (v[:, 1:, :] - v[:, :-1, :]) / dy
         + (w[1:, :, :] - w[:-1, :, :]) / dz) / dt
    f -= f.mean()
    return f.contiguous()


def _release_p(fut):
    if fut.exception() is None:
        fut.result().p = None


def run_stream(server, reqs, realtime=True, progress=None, keep=None):
    """Submit the stream at its arrival times; returns (results, wall_s).
    `keep`: the request ids whose pressure the results hold (None: all); the
    other pressures are released as their requests complete, so a long
    stream holds no more device memory than the server does."""
    import torch
    t0 = time.perf_counter()
    futs = []
    for k, d in enumerate(reqs):
        if realtime:
            delay = t0 + d["t"] - time.perf_counter()
            if delay > 0:
                time.sleep(delay)
        f = device_rhs(d)
        torch.cuda.current_stream().synchronize()
        r = Request(n=d["n"], h=d["h"], bc=d["bc"], f=f, tol=d["tol"], key=d.get("key"),
                    rid=d["rid"], t_submit=t0 + d["t"] if realtime else time.perf_counter())
        fut = server.submit(r)
        if keep is not None and d["rid"] not in keep:
            fut.add_done_callback(_release_p)
        futs.append(fut)
        if progress and (k + 1) % progress == 0:
            print(f"  submitted {k + 1}/{len(reqs)}", file=sys.stderr, flush=True)
    results = [f.result() for f in futs]
    return results, time.perf_counter() - t0


def summarize(results, wall):
    lat = np.array([r.latency_s for r in results])
    ok = sum(1 for r in results if r.status == 0)
    status = collections.Counter(int(r.status) for r in results)
    return {"requests": len(results), "converged": ok, "status_counts": dict(status),
            "wall_s": round(wall, 3),
            "solves_per_s": round(len(results) / wall, 2),
            "latency_ms": {"p50": round(float(np.percentile(lat, 50)) * 1e3, 3),
                           "p99": round(float(np.percentile(lat, 99)) * 1e3, 3),
                           "mean": round(float(lat.mean()) * 1e3, 3)},
            "mean_batch_blocks": round(float(np.mean([r.batch_blocks for r in results])), 3),
            "warm_started": sum(1 for r in results if r.warm)}


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description="serve a seeded Poisson request stream")
    ap.add_argument("--requests", type=int, default=2000)
    ap.add_argument("--rate", type=float, default=200.0)
    ap.add_argument("--seed", type=int, default=4)
    ap.add_argument("--budget-gb", type=float, default=32.0)
    ap.add_argument("--lanes", type=int, default=3)
    ap.add_argument("--no-baseline", action="store_true")
    a = ap.parse_args(argv)
    import torch
    torch.cuda.set_device(0)
    reqs = request_stream(a.requests, a.rate, a.seed)
    out = {"workload": f"{a.requests} requests, Poisson arrivals {a.rate}/s, sizes 32..256 "
                       f"(x16), Neumann/periodic, tol 1e-6/1e-8, 20% warm reuse, seed {a.seed}"}
    with PoissonServer(hbm_budget=int(a.budget_gb * 2**30), lanes=a.lanes) as srv:
        srv.warmup()
        res, wall = run_stream(srv, reqs, progress=500)
        out["served_mixed"] = summarize(res, wall)
        out["served_mixed"]["server"] = dict(srv.stats)
    if not a.no_baseline:
        with PoissonServer(hbm_budget=int(a.budget_gb * 2**30), lanes=1, max_batch_blocks=1,
                           precision="fp64", coarse="smooth", keep_resident=False,
                           maxcy=400) as srv:
            srv.warmup()
            res, wall = run_stream(srv, reqs, progress=500)
            out["sequential_fp64"] = summarize(res, wall)
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
