"""The LES workload's initial field, on the CPU.

u = sin x cos y cos z, v = -cos x sin y cos z, w = 0 sampled on the faces
of the staggered grid (u on x-faces, v on y-faces) is divergence-free under
the face-difference divergence (calmas, src_v70/press.for:50-58) to
round-off: sin(a + h) - sin(a) = 2 sin(h/2) cos(a + h/2) turns both terms
into 2 sin(h/2)/h cos x_c cos y_c cos z_c with opposite signs.  A field
sampled half a cell off (u at cell centres, say) leaves a divergence of
order h, so the check pins the staggering.
"""
import math

import numpy as np


def _div(u, v, w, h):
    return ((np.roll(u, -1, 2) - u) + (np.roll(v, -1, 1) - v) + (np.roll(w, -1, 0) - w)) / h


def test_taylor_green_faces_are_discretely_divergence_free():
    from hydro3d_amd.les import taylor_green
    n = 24
    h = 2.0 * math.pi / n
    u, v, w = taylor_green(n, seed=0, noise=0.0)
    assert np.abs(_div(u, v, w, h)).max() < 1e-12
    assert abs(0.5 * float((u * u + v * v + w * w).mean()) - 0.125) < 1e-12
    # half a cell off in x: no longer divergence-free
    xc = (np.arange(n) + 0.5) * h
    u_c = np.sin(xc)[None, None, :] * np.cos(xc)[None, :, None] * np.cos(xc)[:, None, None]
    assert np.abs(_div(u_c, v, w, h)).max() > 0.1


def test_taylor_green_noise_is_seeded():
    from hydro3d_amd.les import taylor_green
    a = taylor_green(16, seed=3, noise=0.05)
    b = taylor_green(16, seed=3, noise=0.05)
    c = taylor_green(16, seed=4, noise=0.05)
    assert all(np.array_equal(x, y) for x, y in zip(a, b))
    assert not np.array_equal(a[0], c[0])
    clean = taylor_green(16, seed=3, noise=0.0)
    assert 0.04 < float(np.std(a[2] - clean[2])) < 0.06


def test_les_mode_is_checked():
    import pytest
    from hydro3d_amd.les import PeriodicLES
    with pytest.raises(ValueError):
        PeriodicLES(n=8, mode="per_cycle")
