"""Seeded synthetic inputs for the pressure-Poisson workloads.

Array convention: C-ordered numpy arrays of shape (nk, nj, ni), i.e. the
memory of a Fortran (i, j, k) array.  Face velocities: u has shape
(nk, nj, ni+1) where u[..., i+1] is the east face of cell i (the reference's
u(i) in calmas, src_v70/press.for:48-63); v is (nk, nj+1, ni); w is
(nk+1, nj, ni).

Recipe (shared by the CPU oracle, the reference driver and the GPU path):
  * numpy default_rng(seed) (PCG64), standard_normal drawn for u, then v,
    then w, each over its full face array in C order;
  * a periodic axis has face 0 == face n (the wrap), a wall axis has zero
    normal velocity on both wall faces (no penetration), so the divergence
    sums to zero up to rounding;
  * f = div(u)/dt with the 2nd-order stencil of calmas, evaluated as
    ((u_e-u_w)/dx + (v_n-v_s)/dy + (w_t-w_b)/dz)/dt, then f -= mean(f).
"""
from __future__ import annotations

import math

import numpy as np

BC_WALL = 0
BC_PERIODIC = 1


def face_velocity(n, bc, seed, smooth_amp=0.0, h=None):
    """Face velocities for a block of interior size n=(ni,nj,nk)."""
    ni, nj, nk = n
    if h is None:
        h = (1.0 / ni, 1.0 / nj, 1.0 / nk)
    rng = np.random.default_rng(seed)
    u = rng.standard_normal((nk, nj, ni + 1))
    v = rng.standard_normal((nk, nj + 1, ni))
    w = rng.standard_normal((nk + 1, nj, ni))
    if smooth_amp:
        # u += A sin(2 pi x_face) cos(2 pi z_centre) on the x-faces
        x = np.arange(ni + 1) * h[0]
        z = (np.arange(nk) + 0.5) * h[2]
        u += smooth_amp * np.sin(2 * math.pi * x)[None, None, :] * np.cos(2 * math.pi * z)[:, None, None]
    for arr, ax, lo, hi in ((u, 2, bc[0], bc[1]), (v, 1, bc[2], bc[3]), (w, 0, bc[4], bc[5])):
        sl0 = [slice(None)] * 3
        sln = [slice(None)] * 3
        sl0[ax] = 0
        sln[ax] = arr.shape[ax] - 1
        if lo == BC_PERIODIC or hi == BC_PERIODIC:
            arr[tuple(sl0)] = arr[tuple(sln)]
        else:
            arr[tuple(sl0)] = 0.0
            arr[tuple(sln)] = 0.0
    return u, v, w


def extend_faces(u, v, w, bc, seed):
    """The face arrays with one more face at each end of their own axis
    (faces -1..n+1: the second ghost faces calmas' 4th-order branch reads,
    src_v70/press.for:17-37): a periodic axis wraps (face -1 = face n-1,
    face n+1 = face 1); across a wall the outer faces are seeded N(0,1)
    values (the host velocity there is whatever the LES left in it)."""
    rng = np.random.default_rng(seed)
    out = []
    for arr, ax, lo in ((u, 2, bc[0]), (v, 1, bc[2]), (w, 0, bc[4])):
        n = arr.shape[ax] - 1
        if lo == BC_PERIODIC:
            first = np.take(arr, [n - 1], axis=ax)
            last = np.take(arr, [1], axis=ax)
        else:
            shp = list(arr.shape)
            shp[ax] = 1
            first, last = rng.standard_normal(shp), rng.standard_normal(shp)
        out.append(np.ascontiguousarray(np.concatenate([first, arr, last], axis=ax)))
    return tuple(out)


def divergence_rhs(u, v, w, h, dt, remove_mean=True):
    """f = div(u)/dt in the evaluation order of calmas (press.for:50-58)."""
    dx, dy, dz = h
    fact = ((u[:, :, 1:] - u[:, :, :-1]) / dx + (v[:, 1:, :] - v[:, :-1, :]) / dy) + \
        (w[1:, :, :] - w[:-1, :, :]) / dz
    f = fact / dt
    if remove_mean:
        f = f - f.mean()
    return np.ascontiguousarray(f)


# ---------------------------------------------------------------------------
# Named workloads.

def cube(n, bc, nlev, seed, smooth_amp=0.0, dt=1e-3):
    """A single n^3 block on the unit cube (h = 1/n)."""
    h = (1.0 / n, 1.0 / n, 1.0 / n)
    u, v, w = face_velocity((n, n, n), bc, seed, smooth_amp, h)
    f = divergence_rhs(u, v, w, h, dt)
    return dict(n=(n, n, n), h=h, bc=tuple(bc), nlev=nlev, f=f, u=u, v=v, w=w, dt=dt)


# periodic in x and y, Neumann walls in z
BC_PPN = (BC_PERIODIC, BC_PERIODIC, BC_PERIODIC, BC_PERIODIC, BC_WALL, BC_WALL)
BC_WALLS = (BC_WALL,) * 6


def single_64():
    """64^3, dx=dy=dz=1/64, periodic x/y, Neumann z, 4 levels, seed 0."""
    return cube(64, BC_PPN, 4, seed=0)


def single_256():
    """256^3, same boundaries, N(0,1) faces + 5 sin(2 pi x) cos(2 pi z) on u,
    6 levels, seed 1."""
    return cube(256, BC_PPN, 6, seed=1, smooth_amp=5.0)


def uneven_blocks(nblocks=48, seed=2, sizes=(32, 48, 64, 96, 128), nlev=4):
    """Independent blocks with per-axis sizes drawn from `sizes`, Neumann
    walls on all faces, per-block RHS as above with the mean removed."""
    rng = np.random.default_rng(seed)
    dims = rng.choice(np.array(sizes), size=(nblocks, 3))
    out = []
    for b in range(nblocks):
        n = tuple(int(x) for x in dims[b])
        h = (1.0 / 64, 1.0 / 64, 1.0 / 64)
        u, v, w = face_velocity(n, BC_WALLS, seed * 1000 + b, 0.0, h)
        f = divergence_rhs(u, v, w, h, 1e-3)
        out.append(dict(n=n, h=h, bc=BC_WALLS, nlev=nlev, f=f))
    return out


def levels_for(n, cap=6):
    """L = min(cap, min over axes of v2(n)), v2 = exponent of the largest
    power of two dividing n: n/2^(L-1) is then an even integer, as
    src_v70/localparameters.for:202-208 requires (divisible by 2^(L-1))."""
    L = cap
    for x in n:
        v2 = (x & -x).bit_length() - 1
        L = min(L, v2)
    return max(L, 1)
