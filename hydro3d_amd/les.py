"""The projection inside a time-stepping LES (the 192^3 periodic Taylor-Green
workload): every Runge-Kutta stage projects its provisional velocity with
the device multigrid, warm-started from the previous pressure.

Only the pressure solve is this library's.  The momentum terms around it
are a compact stand-in for the reference's host path, written here in torch
on the device:

  * staggered periodic grid, u on x-faces, v on y-faces, w on z-faces;
  * 2nd-order central convection in advective form (differencing = 1);
  * diffusion with nu + nu_t, nu_t = (Cs*delta)^2 |S| (Smagorinsky,
    Cs = 0.1 as src_v70/eddyvis_smag.for:23 hard-codes it);
  * 3-stage SSP Runge-Kutta, in one of two modes:
    - "per_stage": every stage is followed by a projection (u = u* - dt
      grad p with A p = div(u*)/dt);
    - "per_step", the reference's: the intermediate stages only subtract
      the old pressure's gradient (calvel, src_v70/flosol.for:155-182), and
      the pressure is solved once, after the last stage (newsolv_mg,
      flosol.for:211-217).

The reference's own LES (flosol.for and its callees) is not built here, so
the velocity fields -- and therefore the right-hand sides -- are not the
reference's; the solver's behaviour inside a time loop (warm starts, cycles
per solve, solves/s) and the post-projection divergence are what this
measures.
"""
from __future__ import annotations

import math
import time

import numpy as np


def taylor_green(n=192, seed=3, noise=0.05):
    """u = sin x cos y cos z, v = -cos x sin y cos z, w = 0 on [0, 2pi)^3,
    sampled on the faces, plus noise*N(0,1) (numpy default_rng(seed), drawn
    for u, then v, then w, in C order).  Returns (u, v, w) of shape
    (n, n, n), index [k, j, i]; u[k, j, i] is the face west of cell i."""
    h = 2.0 * math.pi / n
    xc = (np.arange(n) + 0.5) * h  # cell centres
    xf = np.arange(n) * h          # west/south/bottom faces
    rng = np.random.default_rng(seed)
    u = (np.sin(xf)[None, None, :] * np.cos(xc)[None, :, None] * np.cos(xc)[:, None, None])
    v = (-np.cos(xc)[None, None, :] * np.sin(xf)[None, :, None] * np.cos(xc)[:, None, None])
    w = np.zeros((n, n, n))
    u = u + noise * rng.standard_normal((n, n, n))
    v = v + noise * rng.standard_normal((n, n, n))
    w = w + noise * rng.standard_normal((n, n, n))
    return u, v, w


def _d(a, axis, h, shift):
    """One-sided difference (a - roll(a, shift)) / h along a periodic axis."""
    import torch
    return (a - torch.roll(a, shifts=shift, dims=axis)) / h


class PeriodicLES:
    """State (u, v, w, p) on the device and the stage operators."""

    def __init__(self, n=192, nu=1e-4, cs=0.1, dt=1.4e-2, seed=3, nlev=6, tol=1e-8,
                 precision="mixed", mode="per_stage"):
        import torch
        from .solver import MultigridSolver
        if mode not in ("per_stage", "per_step"):
            raise ValueError(f"mode {mode!r}: 'per_stage' or 'per_step'")
        self.n, self.nu, self.cs, self.dt, self.mode = n, nu, cs, dt, mode
        self.h = 2.0 * math.pi / n
        u, v, w = taylor_green(n, seed)
        self.u, self.v, self.w = (torch.from_numpy(a).cuda() for a in (u, v, w))
        kw = dict(precision="mixed", coarse="exact") if precision == "mixed" else {}
        self.solver = MultigridSolver([dict(n=(n, n, n), h=(self.h,) * 3, bc=(1,) * 6,
                                            nlev=nlev)], tol=tol, **kw)
        self.p = torch.zeros_like(self.u)
        self.solves = 0
        self.cycles = 0
        self.solve_s = 0.0
        self.max_rmax = 0.0
        self.max_rel_div = 0.0

    def close(self):
        self.solver.close()

    # ---- momentum terms (stand-in)
    def _rhs(self, u, v, w):
        import torch
        h, roll = self.h, torch.roll
        X, Y, Z = 2, 1, 0

        def avg(a, axis, shift):
            return 0.5 * (a + roll(a, shifts=shift, dims=axis))

        # cell-centred velocities and the strain rate for nu_t
        uc, vc, wc = avg(u, X, -1), avg(v, Y, -1), avg(w, Z, -1)

        def cd(a, axis):  # central difference at cell centres
            return (roll(a, -1, axis) - roll(a, 1, axis)) / (2 * h)

        s11, s22, s33 = _d(roll(u, -1, X), X, h, 1), _d(roll(v, -1, Y), Y, h, 1), \
            _d(roll(w, -1, Z), Z, h, 1)
        s12 = 0.5 * (cd(uc, Y) + cd(vc, X))
        s13 = 0.5 * (cd(uc, Z) + cd(wc, X))
        s23 = 0.5 * (cd(vc, Z) + cd(wc, Y))
        smag = torch.sqrt(2.0 * (s11 * s11 + s22 * s22 + s33 * s33 +
                                 2.0 * (s12 * s12 + s13 * s13 + s23 * s23)))
        nut = (self.cs * h) ** 2 * smag

        def lap(a):
            return sum((roll(a, -1, ax) - 2.0 * a + roll(a, 1, ax)) / (h * h) for ax in (0, 1, 2))

        out = []
        for a, ax in ((u, X), (v, Y), (w, Z)):
            # advecting velocities at this face
            ua = u if ax == X else 0.25 * (u + roll(u, -1, X) + roll(u, 1, ax) +
                                           roll(roll(u, -1, X), 1, ax))
            va = v if ax == Y else 0.25 * (v + roll(v, -1, Y) + roll(v, 1, ax) +
                                           roll(roll(v, -1, Y), 1, ax))
            wa = w if ax == Z else 0.25 * (w + roll(w, -1, Z) + roll(w, 1, ax) +
                                           roll(roll(w, -1, Z), 1, ax))
            conv = ua * cd(a, X) + va * cd(a, Y) + wa * cd(a, Z)
            nu_f = self.nu + avg(nut, ax, 1)
            out.append(-conv + nu_f * lap(a))
        return out

    # ---- projection through the device multigrid
    def project(self, us, vs, ws, dt):
        import torch
        h = self.h
        div = (roll_diff(us, 2, h) + roll_diff(vs, 1, h) + roll_diff(ws, 0, h))
        f = (div / dt).contiguous()
        self.solver.set_rhs(0, f - f.mean())
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        st, ncyc, hist = self.solver.solve(raise_on_fail=True)
        torch.cuda.synchronize()
        self.solve_s += time.perf_counter() - t0
        self.solves += 1
        self.cycles += ncyc
        self.solver.get_p(0, out=self.p)
        p = self.p
        u = us - dt * _d(p, 2, h, 1)
        v = vs - dt * _d(p, 1, h, 1)
        w = ws - dt * _d(p, 0, h, 1)
        div_after = roll_diff(u, 2, h) + roll_diff(v, 1, h) + roll_diff(w, 0, h)
        amax = float(div_after.abs().max())
        self.max_rmax = max(self.max_rmax, amax * h ** 3)
        self.max_rel_div = max(self.max_rel_div, amax / max(float(div.abs().max()), 1e-300))
        return u, v, w

    def old_pressure(self, us, vs, ws, dt):
        """An intermediate stage of the reference's mode: u = u* - dt grad p
        with the pressure of the last solve, no solve (calvel)."""
        h, p = self.h, self.p
        return us - dt * _d(p, 2, h, 1), vs - dt * _d(p, 1, h, 1), ws - dt * _d(p, 0, h, 1)

    def step(self):
        """One SSP-RK3 step: a projection after every stage ("per_stage"), or
        the old pressure's gradient after the first two and one projection
        after the last ("per_step").  Stage s builds
        u* = a_s u0 + b_s (u_prev + dt r) with div u0 = div u_prev = 0, so
        div u* = b_s dt div r: projecting with the stage's effective step
        b_s dt (1, 1/4, 2/3) gives every stage the right-hand side div r and
        pressures of one scale, which is what lets each solve warm-start
        from the last.  The projected velocity does not depend on that
        scale (u = u* - b_s dt grad p with A p = div u* / (b_s dt))."""
        dt = self.dt
        u0, v0, w0 = self.u, self.v, self.w
        stage = self.project if self.mode == "per_stage" else self.old_pressure
        r = self._rhs(u0, v0, w0)
        u1, v1, w1 = stage(u0 + dt * r[0], v0 + dt * r[1], w0 + dt * r[2], dt)
        r = self._rhs(u1, v1, w1)
        u2, v2, w2 = stage(0.75 * u0 + 0.25 * (u1 + dt * r[0]),
                           0.75 * v0 + 0.25 * (v1 + dt * r[1]),
                           0.75 * w0 + 0.25 * (w1 + dt * r[2]), 0.25 * dt)
        r = self._rhs(u2, v2, w2)
        self.u, self.v, self.w = self.project(
            u0 / 3.0 + 2.0 / 3.0 * (u2 + dt * r[0]), v0 / 3.0 + 2.0 / 3.0 * (v2 + dt * r[1]),
            w0 / 3.0 + 2.0 / 3.0 * (w2 + dt * r[2]), 2.0 / 3.0 * dt)


def roll_diff(a, axis, h):
    """(a[i+1] - a[i]) / h along a periodic axis: the divergence stencil of
    face values (calmas, src_v70/press.for:50-58)."""
    import torch
    return (torch.roll(a, shifts=-1, dims=axis) - a) / h


def run(steps=50, n=192, **kw):
    """Run `steps` time steps; returns the measured summary."""
    import torch
    les = PeriodicLES(n=n, **kw)
    try:
        les.step()  # first step: cold hierarchy, graph capture
        les.solves = les.cycles = 0
        les.solve_s = 0.0
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            les.step()
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        ke = 0.5 * float((les.u ** 2 + les.v ** 2 + les.w ** 2).mean())
        per = "per stage" if les.mode == "per_stage" else \
            "per step (after the last stage; the earlier stages use the old pressure)"
        return {"workload": f"{n}^3 periodic Taylor-Green + 5% N(0,1) noise (seed 3), nu=1e-4, "
                            f"Smagorinsky Cs=0.1, SSP-RK3, dt={les.dt}, {steps} steps, "
                            f"a warm-started mixed-precision projection (to 1e-8) {per}; "
                            "momentum terms: torch stand-in of the host path",
                "mode": les.mode, "steps": steps, "solves": les.solves,
                "solves_per_s": round(les.solves / les.solve_s, 2),
                "ms_per_solve": round(les.solve_s / les.solves * 1e3, 3),
                "cycles_per_solve": round(les.cycles / les.solves, 2),
                "ms_per_step": round(wall / steps * 1e3, 3),
                "max_post_projection_rmax": les.max_rmax,
                "max_post_projection_div_rel": les.max_rel_div,
                "kinetic_energy": ke}
    finally:
        les.close()
