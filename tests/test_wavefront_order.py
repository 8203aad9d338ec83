"""The evaluation orders the HIP kernels use for the lexicographic smoothers,
restated in Python and checked bitwise against the reference build's
fixtures.  CPU only.

The reference sweeps GS (src_v70/mgsolver.for:407-432) and the three TDMA
passes (:434-549) in place, in loop order.  The kernels run them as
wavefronts instead:

* GS over 8^3 tiles, one tile-diagonal I+J+K = D at a time; inside a tile,
  its own hyperplanes in order (k_gs_tile);
* TDMA over line diagonals (j-lines i+k, i-lines j+k, k-lines i+j); each
  line's right-hand side without its last term first (phase A), then the
  Thomas recurrence, with a(p) and 1/(ap - alo*a(p-1)) taken from one table
  per wall-state pair of the line (k_tdma_diag, k_tdma_tables).

Reads follow the kernels' buffers: not-yet-updated values and ghost images
from the pre-sweep state, updated values from the output.  Equality with the
fixtures shows that the reordering changes no rounding.
"""
import glob
import os

import numpy as np
import pytest

from h3d_util import GOLDEN, P, image_ghosts, load_golden

OPS = sorted(glob.glob(os.path.join(GOLDEN, "ops_*.npz")))
T = 8  # GS tile edge


def _level(z, g):
    bc = tuple(int(x) for x in z["bc"])
    n = tuple(int(x) >> (g - 1) for x in z["n"])
    s = 2.0 ** (g - 1)
    c = tuple(1.0 / ((float(z["h"][a]) * s) * (float(z["h"][a]) * s)) for a in range(3))
    return n, c, bc


def _coef(n, c, bc, i, j, k):
    """coeff (mgsolver.for:185-210): 1/h^2, 0 towards a wall, ap = -1*sum."""
    aw = 0.0 if (bc[0] != P and i == 0) else c[0]
    ae = 0.0 if (bc[1] != P and i == n[0] - 1) else c[0]
    as_ = 0.0 if (bc[2] != P and j == 0) else c[1]
    an = 0.0 if (bc[3] != P and j == n[1] - 1) else c[1]
    ab = 0.0 if (bc[4] != P and k == 0) else c[2]
    at = 0.0 if (bc[5] != P and k == n[2] - 1) else c[2]
    return aw, ae, as_, an, ab, at, -1.0 * (aw + ae + as_ + an + ab + at)


def _image(x, m, per):
    if x < 0:
        return m - 1 if per else 0
    if x >= m:
        return 0 if per else m - 1
    return x


class _Sweep:
    def __init__(self, z, g, scheme):
        self.n, self.c, self.bc = _level(z, g)
        self.phi = z[f"relax{g}_s{scheme}_in_phi"]
        self.rhs = z[f"relax{g}_s{scheme}_in_rhs"]
        self.src = self.phi[1:-1, 1:-1, 1:-1].copy()
        self.dst = np.full_like(self.src, np.nan)
        self.per = (self.bc[0] == P, self.bc[2] == P, self.bc[4] == P)

    def inside(self, x, y, z):
        return 0 <= x < self.n[0] and 0 <= y < self.n[1] and 0 <= z < self.n[2]

    def before(self, x, y, z):
        """pre-sweep value, ghosts as images"""
        n, p = self.n, self.per
        return self.src[_image(z, n[2], p[2]), _image(y, n[1], p[1]), _image(x, n[0], p[0])]

    def result(self):
        out = np.zeros_like(self.phi)
        out[1:-1, 1:-1, 1:-1] = self.dst
        return image_ghosts(out, self.bc)


def gs_tiles(z, g):
    S = _Sweep(z, g, 1)
    ni, nj, nk = S.n
    tl = [(m + T - 1) // T for m in S.n]
    for D in range(sum(tl) - 2):
        for K in range(tl[2]):
            for J in range(tl[1]):
                I = D - K - J
                if not 0 <= I < tl[0]:
                    continue
                o = (I * T, J * T, K * T)
                v = {}
                for lz in range(-1, T + 1):
                    for ly in range(-1, T + 1):
                        for lx in range(-1, T + 1):
                            if (not 0 <= lx < T) + (not 0 <= ly < T) + (not 0 <= lz < T) > 1:
                                continue
                            x, y, zz = o[0] + lx, o[1] + ly, o[2] + lz
                            if not S.inside(x, y, zz):
                                v[lx, ly, lz] = S.before(x, y, zz)
                            elif min(lx, ly, lz) < 0:
                                v[lx, ly, lz] = S.dst[zz, y, x]
                            else:
                                v[lx, ly, lz] = S.src[zz, y, x]
                for h in range(3 * T - 2):
                    for ly in range(T):
                        for lx in range(T):
                            lz = h - lx - ly
                            i, j, k = o[0] + lx, o[1] + ly, o[2] + lz
                            if not (0 <= lz < T and i < ni and j < nj and k < nk):
                                continue
                            aw, ae, as_, an, ab, at, ap = _coef(S.n, S.c, S.bc, i, j, k)
                            sm = (aw * v[lx - 1, ly, lz] + ae * v[lx + 1, ly, lz] +
                                  as_ * v[lx, ly - 1, lz] + an * v[lx, ly + 1, lz] +
                                  ab * v[lx, ly, lz - 1] + at * v[lx, ly, lz + 1])
                            val = (S.rhs[k, j, i] - sm) / ap
                            v[lx, ly, lz] = val
                            S.dst[k, j, i] = val
    return S.result()


def _cell(pas, u, v, p):
    return (u, p, v) if pas == 0 else ((p, u, v) if pas == 1 else (u, v, p))


def _wall_state(c, m, per):
    return 0 if per else (1 if c == 0 else 0) | (2 if c == m - 1 else 0)


def _wall_rep(s, m, per):
    if per:
        return 0 if s == 0 else -1
    if s == 0:
        return 1 if m >= 3 else -1
    if s in (1, 2):
        return (0 if s == 1 else m - 1) if m >= 2 else -1
    return 0 if m == 1 else -1


def tdma_diagonals(z, g):
    S = _Sweep(z, g, 2)
    ni, nj, nk = S.n
    for pas in range(3):
        nu = nj if pas == 1 else ni
        nv = nj if pas == 2 else nk
        ln = (nj, ni, nk)[pas]
        pu = S.per[1] if pas == 1 else S.per[0]
        pv = S.per[1] if pas == 2 else S.per[2]
        lo_hi = ((2, 3), (0, 1), (4, 5))[pas]
        # k_tdma_tables: a(p), 1/(ap - alo*a(p-1)) per wall-state pair
        tab = {}
        for var in range(16):
            ru, rv = _wall_rep(var & 3, nu, pu), _wall_rep(var >> 2, nv, pv)
            if ru < 0 or rv < 0:
                continue
            a, rows = 0.0, []
            for p in range(ln):
                cf = _coef(S.n, S.c, S.bc, *_cell(pas, ru, rv, p))
                apr = 1.0 / (cf[6] - cf[lo_hi[0]] * a)
                a = cf[lo_hi[1]] * apr
                rows.append((a, apr))
            tab[var] = rows

        def cross(x, y, zz, updated):
            if not S.inside(x, y, zz):
                return S.before(x, y, zz)
            return S.dst[zz, y, x] if (updated or pas > 0) else S.src[zz, y, x]

        for d in range(nu + nv - 1):
            for u in range(max(0, d - nv + 1), min(nu, d + 1)):
                v = d - u
                rs = []
                for p in range(ln):  # phase A
                    i, j, k = _cell(pas, u, v, p)
                    aw, ae, as_, an, ab, at, _ = _coef(S.n, S.c, S.bc, i, j, k)
                    r = S.rhs[k, j, i]
                    if pas == 0:
                        r = (r - ae * cross(i + 1, j, k, False) - aw * cross(i - 1, j, k, True) -
                             at * cross(i, j, k + 1, False) - ab * cross(i, j, k - 1, True))
                    elif pas == 1:
                        r = (r - an * cross(i, j + 1, k, False) - as_ * cross(i, j - 1, k, True) -
                             at * cross(i, j, k + 1, False) - ab * cross(i, j, k - 1, True))
                    else:
                        r = (r - an * cross(i, j + 1, k, False) - as_ * cross(i, j - 1, k, True) -
                             ae * cross(i + 1, j, k, False) - aw * cross(i - 1, j, k, True))
                    rs.append(r)
                rows = tab[_wall_state(u, nu, pu) | (_wall_state(v, nv, pv) << 2)]
                cp = S.before(*_cell(pas, u, v, -1))
                cs = []
                for p in range(ln):  # phase B, forward
                    alo = _coef(S.n, S.c, S.bc, *_cell(pas, u, v, p))[lo_hi[0]]
                    cp = (rs[p] - alo * cp) * rows[p][1]
                    cs.append(cp)
                nx = S.before(*_cell(pas, u, v, ln))
                for p in range(ln - 1, -1, -1):  # back substitution
                    nx = cs[p] - rows[p][0] * nx
                    i, j, k = _cell(pas, u, v, p)
                    S.dst[k, j, i] = nx
    return S.result()


@pytest.mark.parametrize("path", OPS, ids=[os.path.basename(p) for p in OPS])
def test_gs_tile_wavefront_is_the_reference_order(path):
    z = load_golden(os.path.basename(path))
    for g in range(1, int(z["nlev"]) + 1):
        assert np.array_equal(gs_tiles(z, g), z[f"relax{g}_s1_out"]), f"level {g}"


@pytest.mark.parametrize("path", OPS, ids=[os.path.basename(p) for p in OPS])
def test_tdma_line_diagonals_are_the_reference_order(path):
    z = load_golden(os.path.basename(path))
    for g in range(1, int(z["nlev"]) + 1):
        assert np.array_equal(tdma_diagonals(z, g), z[f"relax{g}_s2_out"]), f"level {g}"


def test_wavefront_order_check_detects_a_wrong_order():
    """A wrong order (every neighbour read as before the sweep: Jacobi) fails
    the same comparison, so the check tells evaluation orders apart."""
    z = load_golden("ops_ppw8.npz")
    S = _Sweep(z, 1, 1)
    ni, nj, nk = S.n
    for k in range(nk):
        for j in range(nj):
            for i in range(ni):
                aw, ae, as_, an, ab, at, ap = _coef(S.n, S.c, S.bc, i, j, k)
                sm = (aw * S.before(i - 1, j, k) + ae * S.before(i + 1, j, k) +
                      as_ * S.before(i, j - 1, k) + an * S.before(i, j + 1, k) +
                      ab * S.before(i, j, k - 1) + at * S.before(i, j, k + 1))
                S.dst[k, j, i] = (S.rhs[k, j, i] - sm) / ap
    assert not np.array_equal(S.result(), z["relax1_s1_out"])
