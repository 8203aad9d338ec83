"""Shared helpers of the test suite (index identities and spectral solves in
numpy; nothing here calls the product)."""
from __future__ import annotations

import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
P, W = 1, 0


def load_golden(name):
    with np.load(os.path.join(GOLDEN, name), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def image_ghosts(phi, bc):
    """Ghost layer of a padded (nk+2, nj+2, ni+2) array set to the image of
    the interior (mirror at walls, wrap at periodic faces)."""
    phi = phi.copy()
    for ax, (lo, hi) in zip((2, 1, 0), ((bc[0], bc[1]), (bc[2], bc[3]), (bc[4], bc[5]))):
        n = phi.shape[ax] - 2

        def sl(x):
            return tuple(x if a == ax else slice(None) for a in range(3))

        if lo == P:
            phi[sl(0)] = phi[sl(n)]
            phi[sl(n + 1)] = phi[sl(1)]
        else:
            phi[sl(0)] = phi[sl(1)]
            phi[sl(n + 1)] = phi[sl(n)]
    return phi


def op1d(n, h, periodic):
    """Dense 1-D operator of the 7-point stencil along one axis: 1/h^2 to
    each neighbour, no coupling through a wall (zero wall coefficient)."""
    A = np.zeros((n, n))
    c = 1.0 / (h * h)
    for i in range(n):
        for d in (-1, 1):
            j = i + d
            if 0 <= j < n:
                A[i, j] += c
                A[i, i] -= c
            elif periodic:
                A[i, j % n] += c
                A[i, i] -= c
    return A


def spectral_solve(f, h, bc):
    """Solve A p = f (f: (nk,nj,ni)) by eigendecomposition of the three 1-D
    operators (numpy.linalg.eigh), dropping the null-space mode; mean-free p."""
    nk, nj, ni = f.shape
    mats = [op1d(ni, h[0], bc[0] == P), op1d(nj, h[1], bc[2] == P), op1d(nk, h[2], bc[4] == P)]
    ev = [np.linalg.eigh(A) for A in mats]
    (lx, Qx), (ly, Qy), (lz, Qz) = ev
    fh = np.einsum("ia,jb,kc,kji->cba", Qx, Qy, Qz, f, optimize=True)
    lam = lz[:, None, None] + ly[None, :, None] + lx[None, None, :]
    small = np.abs(lam) < 1e-9 * np.abs(lam).max()
    ph = np.where(small, 0.0, fh / np.where(small, 1.0, lam))
    p = np.einsum("ia,jb,kc,cba->kji", Qx, Qy, Qz, ph, optimize=True)
    return p - p.mean()


def apply_np(p_pad, h, bc):
    """A p on the interior of a padded p whose ghosts are images (numpy,
    vectorised; independent of the oracle and of the product)."""
    c = [1.0 / (x * x) for x in h]
    q = p_pad
    mid = q[1:-1, 1:-1, 1:-1]
    out = np.zeros_like(mid)
    for ax, (lo, cc) in zip((2, 1, 0), ((bc[0], c[0]), (bc[2], c[1]), (bc[4], c[2]))):
        sl_m = [slice(1, -1)] * 3
        sl_p = [slice(1, -1)] * 3
        sl_m[ax] = slice(0, -2)
        sl_p[ax] = slice(2, None)
        lap = q[tuple(sl_m)] - mid + q[tuple(sl_p)] - mid
        if lo != P:
            # wall: drop the flux through the first/last face
            n = mid.shape[ax]
            first = [slice(None)] * 3
            last = [slice(None)] * 3
            first[ax] = 0
            last[ax] = n - 1
            lapf = lap[tuple(first)] - (q[tuple(sl_m)][tuple(first)] - mid[tuple(first)])
            lapl = lap[tuple(last)] - (q[tuple(sl_p)][tuple(last)] - mid[tuple(last)])
            lap[tuple(first)] = lapf
            lap[tuple(last)] = lapl
        out += cc * lap
    return out


def mean_free(a):
    return a - a.mean()


def rel_l2(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))
