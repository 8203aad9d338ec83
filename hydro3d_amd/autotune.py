"""Autotuner of the device multigrid's launch shapes and schedule.

These knobs change only how the work is launched, never the results (every
cell is computed by the same arithmetic whatever the choice):

* smoother tile shape per level -- lanes per row segment
  (h3dmg_set_lanes: the row's count, a half or a quarter; the tile's x/y
  aspect) crossed with the march depth, the planes each smoother workgroup walks
  (h3dmg_set_march: 1, 2 or 4): deeper tiles reuse more of the register
  window and keep more loads in flight per wave, shallower ones need fewer
  registers (more waves per SIMD) and spread a small level over more
  workgroups;
* tail level -- the level from which one workgroup per block runs the rest
  of the V-cycle in a single launch (h3dmg_set_tail): below some size a level
  is launch-latency-bound and on-chip barriers beat full-chip launches;
* fused sweeps -- the largest level (cells per block) whose red-black sweeps
  run as one launch each instead of two (h3dmg_set_fused, 0 = none).

`autotune(solver)` measures the candidates on the solver's own blocks and
right-hand side, applies the fastest and returns the choice; results are
cached in JSON keyed by the device, the precision settings and the block
set, so a later `autotune` or `apply_cached` on the same configuration costs
nothing.

    python -m hydro3d_amd.autotune --workload single_256 [--cache PATH]
"""
from __future__ import annotations

import argparse
import json
import os
import time

from . import _lib as L

DEFAULT_CACHE = os.environ.get("H3DMG_TUNE_CACHE",
                               os.path.join(os.path.expanduser("~"), ".cache", "h3dmg",
                                            "autotune.json"))
MARCH_CANDIDATES = (1, 2, 4)
# smoother lanes per row segment: the whole row's count divided by these
LANES_DIVISORS = (1, 2, 4)
TAIL_MAX_CELLS = 64 ** 3  # levels larger than this never go to the one-workgroup tail


def config_key(solver, device: str | None = None) -> str:
    """Cache key: device, arithmetic settings and the block set."""
    p = solver.params
    blocks = sorted((tuple(b["n"]), tuple(int(x) for x in b["bc"]), b["nlev"])
                    for b in solver.blocks)
    key = {"device": device or _device_name(), "precision": int(p.precision),
           "coarse": int(p.coarse), "nu": [int(p.nu1), int(p.nu2)], "blocks": blocks}
    if int(p.smoother) != L.SMOOTH_RBGS:
        key["smoother"] = int(p.smoother)
    return json.dumps(key, sort_keys=True)


def _device_name() -> str:
    """The device's ISA and CU count (e.g. "gfx950 256CU"): the marketing
    name is not reliable (boxes without the device-name table report a
    generic one)."""
    try:
        import torch
        if torch.cuda.is_available():
            pr = torch.cuda.get_device_properties(0)
            arch = getattr(pr, "gcnArchName", "") or torch.cuda.get_device_name(0)
            return f"{arch.split(':')[0]} {pr.multi_processor_count}CU"
    except Exception:  # pragma: no cover - torch absent
        pass
    return "unknown"


def load_cache(path: str = DEFAULT_CACHE) -> dict:
    try:
        with open(path) as fh:
            return json.load(fh)
    except (OSError, ValueError):
        return {}


def save_cache(cache: dict, path: str = DEFAULT_CACHE) -> None:
    d = os.path.dirname(path)
    if d:
        os.makedirs(d, exist_ok=True)
    tmp = path + ".tmp"
    with open(tmp, "w") as fh:
        json.dump(cache, fh, indent=1, sort_keys=True)
    os.replace(tmp, path)


FUSED_CANDIDATES = (0, 32 ** 3, 64 ** 3)


def apply(solver, choice: dict) -> None:
    """Apply a tuning choice {"march": {level: planes}, "tail": level,
    "fused": max_cells}."""
    if "fused" in choice:
        solver.set_fused(int(choice["fused"]))
    for g, lanes in choice.get("lanes", {}).items():
        solver.set_lanes(int(g), int(lanes))
    for g, kb in choice.get("march", {}).items():
        solver.set_march(int(g), int(kb))
    if "tail" in choice:
        solver.set_tail(int(choice["tail"]))


def apply_cached(solver, path: str = DEFAULT_CACHE) -> dict | None:
    choice = load_cache(path).get(config_key(solver))
    if choice is not None:
        apply(solver, choice)
    return choice


def _level_cells(solver, g: int) -> int:
    cells = 0
    for b in solver.blocks:
        if b["nlev"] >= g:
            n = [x >> (g - 1) for x in b["n"]]
            cells = max(cells, n[0] * n[1] * n[2])
    return cells


def _median(xs):
    xs = sorted(xs)
    return xs[len(xs) // 2]


def _time_solve(solver, reps: int) -> float:
    """Median wall time of a solve from p = 0 (device-resident data)."""
    import torch
    zeros = [torch.zeros(tuple(reversed(b["n"])), dtype=torch.float64, device="cuda")
             for b in solver.blocks]
    ts = []
    for _ in range(reps + 1):
        for q, z in enumerate(zeros):
            solver.set_p(q, z)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        solver.solve()
        torch.cuda.synchronize()
        ts.append(time.perf_counter() - t0)
    return _median(ts[1:])


def _time_relax(solver, g: int, reps: int, nsweep: int = 4) -> float:
    ts = []
    for _ in range(reps + 1):
        t0 = time.perf_counter()
        solver.dbg_relax(g, nsweep)  # synchronous
        ts.append(time.perf_counter() - t0)
    return _median(ts[1:])


def autotune(solver, cache_path: str | None = DEFAULT_CACHE, reps: int = 5,
             force: bool = False, verbose: bool = False) -> dict:
    """Measure, apply and cache the fastest march depths and tail level.

    The solver's right-hand sides must be set.  Its pressures are saved and
    restored around the measurement solves.  The knobs belong to the
    red-black smoother's launches: a lexicographic GS or line-TDMA handle
    has none, and gets an empty choice."""
    import torch
    if int(solver.params.smoother) != L.SMOOTH_RBGS:
        return {}
    key = config_key(solver)
    cache = load_cache(cache_path) if cache_path else {}
    if not force and key in cache:
        apply(solver, cache[key])
        return cache[key]
    saved = [solver.get_p(q, out=torch.empty(tuple(reversed(b["n"])), dtype=torch.float64,
                                              device="cuda"))
             for q, b in enumerate(solver.blocks)]
    Lmax = solver.nlevels_max
    choice = {"march": {}, "lanes": {}, "tail": 0}
    log = {}
    # 1. smoother tile shape per level: lanes per row segment (the tile's
    #    x/y aspect) x march depth, timed on the level's sweeps
    for g in range(1, Lmax + 1):
        nk = max((b["n"][2] >> (g - 1)) for b in solver.blocks if b["nlev"] >= g)
        solver.set_lanes(g, 0)
        solver.set_march(g, 0)
        natural = solver.get_lanes(g)
        best, tbest = (0, 0), _time_relax(solver, g, reps)
        log[f"level{g}_auto"] = tbest
        for lanes in LANES_DIVISORS:
            lw = natural // lanes
            if lw < 1 or (lanes > 1 and lw == natural):
                continue
            solver.set_lanes(g, 0 if lanes == 1 else lw)
            for kb in (0,) + MARCH_CANDIDATES:
                if kb > nk or (lanes == 1 and kb == 0):
                    continue
                solver.set_march(g, kb)
                t = _time_relax(solver, g, reps)
                log[f"level{g}_lanes{lw}_march{kb}"] = t
                if t < tbest * 0.98:
                    best, tbest = (0 if lanes == 1 else lw, kb), t
        solver.set_lanes(g, best[0])
        solver.set_march(g, best[1])
        choice["lanes"][str(g)] = best[0]
        choice["march"][str(g)] = best[1]
    # 2. tail level on whole solves
    cands = [0] + [g for g in range(1, Lmax + 1) if _level_cells(solver, g) <= TAIL_MAX_CELLS]
    cands.append(Lmax + 1)  # off
    best, tbest = None, float("inf")
    for lt in cands:
        try:
            solver.set_tail(lt)
        except L.H3dError:  # the remaining levels do not fit in LDS
            continue
        t = _time_solve(solver, reps)
        log[f"tail{lt}"] = t
        if t < tbest * (0.98 if best is not None else 1.0):
            best, tbest = lt, t
    solver.set_tail(best)
    choice["tail"] = best
    # 3. fused sweeps on whole solves
    fbest = 0
    for fz in FUSED_CANDIDATES[1:]:
        solver.set_fused(fz)
        t = _time_solve(solver, reps)
        log[f"fused{fz}"] = t
        if t < tbest * 0.98:
            fbest, tbest = fz, t
    solver.set_fused(fbest)
    choice["fused"] = fbest
    choice["solve_ms"] = tbest * 1e3
    # the space searched: every candidate's median time (ms; level* =
    # four sweeps of that level, tail* / fused* = whole solves)
    choice["searched"] = {k: round(v * 1e3, 4) for k, v in log.items()}
    for q, p in enumerate(saved):
        solver.set_p(q, p)
    torch.cuda.synchronize()
    if verbose:
        for k, v in log.items():
            print(f"  {k:>14s}: {v * 1e3:8.3f} ms")
    if cache_path:
        cache = load_cache(cache_path)
        cache[key] = choice
        save_cache(cache, cache_path)
    return choice


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description="tune launch shapes for a workload")
    ap.add_argument("--workload", default="single_256",
                    choices=["single_64", "single_256", "uneven_blocks"])
    ap.add_argument("--precision", default="mixed", choices=["fp64", "mixed"])
    ap.add_argument("--cache", default=DEFAULT_CACHE)
    ap.add_argument("--reps", type=int, default=5)
    a = ap.parse_args(argv)
    import torch
    from . import MultigridSolver, configs
    torch.cuda.set_device(0)
    if a.workload == "uneven_blocks":
        cs = configs.uneven_blocks()
    else:
        cs = [getattr(configs, a.workload)()]
    specs = [dict(n=c["n"], h=c["h"], bc=c["bc"], nlev=c["nlev"]) for c in cs]
    coarse = "exact" if a.precision == "mixed" else "smooth"
    with MultigridSolver(specs, precision=a.precision, coarse=coarse) as S:
        for q, c in enumerate(cs):
            S.set_rhs(q, torch.from_numpy(c["f"]).cuda())
        choice = autotune(S, a.cache, reps=a.reps, force=True, verbose=True)
    print(json.dumps(choice))
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
