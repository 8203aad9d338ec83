"""Serving: continuous batching of Poisson solves from concurrent simulations.

Requests (one block each: size, spacing, boundaries, right-hand side,
tolerance, optional residency key) arrive at any time.  The server

* admits queued requests by HBM footprint (hierarchies of every admitted or
  resident request must fit the budget; idle resident hierarchies are evicted
  least-recently-used first),
* batches the requests that are waiting when a worker slot frees up into ONE
  multi-block handle -- one launch per level and operator covers all of them,
  each block with its own tolerance (h3dmg_set_block_tol) and its own
  retirement,
* runs up to `max_batches` handles at once, each on its own thread and HIP
  stream, so that work admitted while others are on the GPU overlaps with
  them (continuous batching, not a queue),
* keeps every finished hierarchy and its pressure resident under the
  request's key: a later request with the same key reuses it, warm-started
  from the previous pressure (h3dmg_set_block_enabled restricts the solve to
  the requested blocks of a resident handle).

Blocks of one handle never interact (per-block norms, tolerances and
retirement), so each result is the one the request gets alone.

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


# ---- footprint ------------------------------------------------------------
def _align(x, a=256):
    return (x + a - 1) // a * a


def footprint_bytes(blocks, precision="mixed", coarse="exact") -> int:
    """Device bytes a handle over `blocks` allocates (the arena layout of
    h3dmg_create plus its separately allocated tile tables, rounded up)."""
    mixed = precision == "mixed"
    exact = coarse == "exact"
    total = 0
    stage = 0
    for b in blocks:
        ni, nj, nk = b["n"]
        L = b["nlev"]
        for g in range(1, L + 1):
            n = (ni >> (g - 1), nj >> (g - 1), nk >> (g - 1))
            H = (((n[0] + 1) // 2) + 1) & ~1
            w = 4 if (mixed and g < L) else 8
            total += 6 * _align(H * n[1] * n[2] * w)
            # tile tables (smoother, restriction, prolongation, zero fill)
            total += 24 * max(1, (H * n[1] * n[2]) // 512) * 4
        H = (((ni + 1) // 2) + 1) & ~1
        total += 4 * _align(H * nj * nk * 8)
        if exact:
            c = (ni >> (L - 1), nj >> (L - 1), nk >> (L - 1))
            total += sum(_align(x * x * 8) + _align(x * 8) for x in c) + 2 * _align(c[0] * c[1] * c[2] * 8)
        stage = max(stage, (ni + 2) * (nj + 2) * (nk + 2) * 8)
    return int((total + stage) * 1.05) + (1 << 20)


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
    status: int
    cycles: int
    rel_residual: float
    latency_s: float
    batch_blocks: int
    warm: bool


@dataclass
class _Resident:
    solver: object
    keys: list                # key of each block (None: not reusable)
    nbytes: int
    busy: bool = False
    last_used: float = field(default_factory=time.perf_counter)
    specs: list = None        # block specs (known before the handle is built)


class PoissonServer:
    """Continuous-batching server over the device multigrid.

    solver_factory(specs, **kw) builds a handle (default: MultigridSolver);
    tests substitute a host model of it."""

    def __init__(self, hbm_budget=64 << 30, max_batches=4, max_batch_blocks=16,
                 precision="mixed", coarse="exact", nu1=2, nu2=2, maxcy=60,
                 solver_factory=None, keep_resident=True, use_graph=False):
        if solver_factory is None:
            from .solver import MultigridSolver as solver_factory  # noqa: N813
        self._factory = solver_factory
        self.precision, self.coarse = precision, coarse
        # Handles run on several host threads; by default they launch eagerly
        # (each on its own stream), since a handle's first solve would
        # otherwise capture its graph while other threads create handles.
        self._kw = dict(nu1=nu1, nu2=nu2, maxcy=maxcy, precision=precision, coarse=coarse,
                        use_graph=use_graph)
        self.budget = int(hbm_budget)
        self.max_batches = max_batches
        self.max_batch_blocks = max_batch_blocks
        self.keep_resident = keep_resident
        self._lock = threading.Condition()
        self._pending = collections.deque()          # (Request, Future) for new hierarchies
        self._reuse = collections.deque()            # (Request, Future, _Resident, block)
        self._resident = []                          # list of _Resident
        self._by_key = {}                            # key -> (_Resident, block)
        self._running = 0
        self._closed = False
        self.stats = collections.Counter()
        self._thread = threading.Thread(target=self._dispatch, name="h3dmg-dispatch", daemon=True)
        self._thread.start()

    # -- public
    def submit(self, req: Request) -> Future:
        fut = Future()
        req.t_submit = req.t_submit or time.perf_counter()
        with self._lock:
            if self._closed:
                raise RuntimeError("server closed")
            hit = self._by_key.get(req.key) if req.key is not None else None
            if hit is not None:
                res, b = hit
                sp = res.specs[b]
                if tuple(sp["n"]) == tuple(req.n) and tuple(sp["bc"]) == tuple(req.bc):
                    self._reuse.append((req, fut, res, b))
                    self._lock.notify_all()
                    return fut
            self._pending.append((req, fut))
            self._lock.notify_all()
        return fut

    def warmup(self, n=(32, 32, 32)):
        """Serve one small request and forget it: the kernels' code objects
        and the device context are loaded before traffic arrives, instead of
        by the first requests (whose handles are built concurrently).  The
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
        self._thread.join()
        for r in self._resident:
            if r.solver is not None:
                r.solver.close()
        self._resident.clear()
        self._by_key.clear()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    @property
    def resident_bytes(self):
        with self._lock:
            return sum(r.nbytes for r in self._resident)

    # -- dispatcher
    def _used(self):
        return sum(r.nbytes for r in self._resident)

    def _evict_for(self, need):
        """Evict idle resident handles (LRU) until `need` more bytes fit."""
        idle = sorted((r for r in self._resident if not r.busy), key=lambda r: r.last_used)
        while self._used() + need > self.budget and idle:
            r = idle.pop(0)
            if any(x[2] is r for x in self._reuse):
                continue
            self._resident.remove(r)
            for k in r.keys:
                if k is not None and self._by_key.get(k, (None,))[0] is r:
                    del self._by_key[k]
            r.solver.close()
            self.stats["evicted"] += 1
        return self._used() + need <= self.budget

    def _dispatch(self):
        while True:
            with self._lock:
                while not self._closed and (self._running >= self.max_batches or
                                            not (self._pending or self._ready_reuse())):
                    self._lock.wait()
                if self._closed and not self._pending and not self._reuse and self._running == 0:
                    return
                if self._closed and self._running >= self.max_batches:
                    self._lock.wait(0.01)
                    continue
                job = self._next_job()
                if job is None:
                    if self._closed and not self._running:
                        # nothing admissible will ever fit: fail what is left
                        while self._pending:
                            req, fut = self._pending.popleft()
                            fut.set_exception(MemoryError("request does not fit the HBM budget"))
                        return
                    self._lock.wait(0.05)
                    continue
                self._running += 1
            threading.Thread(target=self._run_job, args=job, daemon=True).start()

    def _ready_reuse(self):
        return any(not r.busy for (_, _, r, _) in self._reuse)

    def _next_job(self):
        # 1. warm-start requests on idle resident handles (all requests for
        #    the same handle go together)
        for (req, fut, res, b) in list(self._reuse):
            if res.busy:
                continue
            group = [(q, f, bb) for (q, f, r, bb) in self._reuse if r is res]
            seen, batch = set(), []
            for q, f, bb in group:
                if bb in seen:
                    continue  # a second request for the same block waits its turn
                seen.add(bb)
                batch.append((q, f, bb))
            for q, f, bb in batch:
                self._reuse.remove(next(x for x in self._reuse if x[0] is q))
            res.busy = True
            return (res, batch, True)
        # 2. a new batch from the queue, admitted by footprint
        if not self._pending:
            return None
        batch, specs = [], []
        while self._pending and len(batch) < self.max_batch_blocks:
            req, fut = self._pending[0]
            trial = specs + [req.spec()]
            need = footprint_bytes(trial, self.precision, self.coarse)
            if need > self.budget:
                if not batch:  # alone it can never fit
                    self._pending.popleft()
                    fut.set_exception(MemoryError(f"request needs {need} bytes, budget {self.budget}"))
                    continue
                break
            if not self._evict_for(need):
                break
            self._pending.popleft()
            batch.append((req, fut))
            specs = trial
        if not batch:
            return None
        # The footprint is reserved here, under the lock; the handle itself is
        # built by the worker (device allocations, tables and uploads take
        # milliseconds, and building here would hold every other dispatch).
        res = _Resident(solver=None, keys=[r.key for r, _ in batch],
                        nbytes=footprint_bytes(specs, self.precision, self.coarse), busy=True,
                        specs=specs)
        self._resident.append(res)
        for b, (r, _) in enumerate(batch):
            if r.key is not None:
                self._by_key[r.key] = (res, b)
        self.stats["batches"] += 1
        self.stats["batched_blocks"] += len(batch)
        return (res, [(r, f, b) for b, (r, f) in enumerate(batch)], False,
                min(r.tol for r, _ in batch))

    # -- worker
    def _drop(self, res):
        """Forget resident `res` (lock held)."""
        if res in self._resident:
            self._resident.remove(res)
        for k in res.keys:
            if k is not None and self._by_key.get(k, (None,))[0] is res:
                del self._by_key[k]

    def _run_job(self, res, batch, warm, tol=None):
        if res.solver is None:
            try:
                res.solver = self._factory(res.specs, tol=tol, **self._kw)
            except Exception as e:  # noqa: BLE001 - device refused: fail this batch
                with self._lock:
                    self._drop(res)
                    # requests queued meanwhile for this handle's keys fail too
                    for x in [x for x in self._reuse if x[2] is res]:
                        self._reuse.remove(x)
                        x[1].set_exception(e)
                    for _, fut, _ in batch:
                        fut.set_exception(e)
                    self._running -= 1
                    self._lock.notify_all()
                return
        S = res.solver
        try:
            blocks = {b for (_, _, b) in batch}
            for b in range(len(S.blocks)):
                S.set_block_enabled(b, b in blocks)
            for req, _, b in batch:
                S.set_rhs(b, req.f)
                S.set_block_tol(b, req.tol)
            st, _, _ = S.solve()
            cyc, rel = S.block_stats()
            done = time.perf_counter()
            for req, fut, b in batch:
                p = S.result_p(b)
                fut.set_result(Result(rid=req.rid, p=p, status=block_status(st, rel[b], req.tol),
                                      cycles=cyc[b],
                                      rel_residual=rel[b], latency_s=done - req.t_submit,
                                      batch_blocks=len(batch), warm=warm))
            self.stats["solves"] += len(batch)
            self.stats["warm"] += len(batch) if warm else 0
        except Exception as e:  # noqa: BLE001
            for _, fut, _ in batch:
                if not fut.done():
                    fut.set_exception(e)
        finally:
            with self._lock:
                res.busy = False
                res.last_used = time.perf_counter()
                if not self.keep_resident and not any(x[2] is res for x in self._reuse):
                    self._drop(res)
                    S.close()
                self._running -= 1
                self._lock.notify_all()


def block_status(batch_status, rel, tol):
    """Status of one block of a multi-block solve: the handle's status
    speaks for the whole batch (NOT_CONVERGED when any block reached maxcy),
    so each block is judged by its own final relative residual; a diverged
    or failed solve keeps the batch's status."""
    if batch_status not in (0, 5):
        return int(batch_status)
    return 0 if math.isfinite(rel) and rel <= tol else 5


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
    ap.add_argument("--budget-gb", type=float, default=64.0)
    ap.add_argument("--max-batches", type=int, default=4)
    ap.add_argument("--no-baseline", action="store_true")
    a = ap.parse_args(argv)
    import torch
    torch.cuda.set_device(0)
    reqs = request_stream(a.requests, a.rate, a.seed)
    out = {"workload": f"{a.requests} requests, Poisson arrivals {a.rate}/s, sizes 32..256 "
                       f"(x16), Neumann/periodic, tol 1e-6/1e-8, 20% warm reuse, seed {a.seed}"}
    with PoissonServer(hbm_budget=int(a.budget_gb * 2**30), max_batches=a.max_batches) as srv:
        srv.warmup()
        res, wall = run_stream(srv, reqs, progress=500)
        out["served_mixed"] = summarize(res, wall)
        out["served_mixed"]["server"] = dict(srv.stats)
    if not a.no_baseline:
        with PoissonServer(hbm_budget=int(a.budget_gb * 2**30), max_batches=1, max_batch_blocks=1,
                           precision="fp64", coarse="smooth", keep_resident=False,
                           maxcy=400) as srv:
            srv.warmup()
            res, wall = run_stream(srv, reqs, progress=500)
            out["sequential_fp64"] = summarize(res, wall)
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
