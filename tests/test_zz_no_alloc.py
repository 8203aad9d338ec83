"""h3dmg_project_device_noalloc against h3dmg_project_device, through the
C-ABI, on two calls in a row (a 32^3 block, then a smaller 16x8x16 block of
another handle): u, v, w, su bitwise and rmax equal.  Needs an MI355X."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).cuda()


def _vp(t):
    return ctypes.c_void_p(t.data_ptr())


def _faces(n, seed):
    ni, nj, nk = n
    rng = np.random.default_rng(seed)
    return (rng.standard_normal((nk, nj, ni + 1)), rng.standard_normal((nk, nj + 1, ni)),
            rng.standard_normal((nk + 1, nj, ni)), rng.standard_normal((nk, nj, ni)))


@pytest.mark.parametrize("bc", [(1, 1, 1, 1, 0, 0), (0, 0, 0, 0, 0, 0)])
def test_gpu_project_noalloc_equals_project(gpu_lib, bc):
    import torch
    from hydro3d_amd import MultigridSolver
    from hydro3d_amd import _lib as L
    lib = L.load()
    for n, seed in (((32, 32, 32), 5), ((16, 8, 16), 6)):
        print(f"case: noalloc projection n={n} bc={bc}", flush=True)
        spec = dict(n=n, h=tuple(1.0 / x for x in n), bc=bc, nlev=3)
        us, vs, ws, p = _faces(n, seed)
        with MultigridSolver([spec], dt=1e-3) as S:
            S.set_p(0, p)
            u0, v0, w0, su0, rmax0 = S.project(0, _dev(us), _dev(vs), _dev(ws), alfapr=0.7)
            d_us, d_vs, d_ws = _dev(us), _dev(vs), _dev(ws)
            u1, v1, w1 = (torch.full_like(t, np.nan) for t in (d_us, d_vs, d_ws))
            su1 = torch.full(n[::-1], np.nan, dtype=torch.float64, device="cuda")
            rmax1 = torch.full((1,), np.nan, dtype=torch.float64, device="cuda")
            torch.cuda.synchronize()
            L.check(lib.h3dmg_project_device_noalloc(
                S.h, 0, _vp(d_us), _vp(d_vs), _vp(d_ws), _vp(u1), _vp(v1), _vp(w1), _vp(su1),
                0.7, _vp(rmax1)))
            torch.cuda.synchronize()
        assert torch.equal(u0, u1) and torch.equal(v0, v1) and torch.equal(w0, w1)
        assert torch.equal(su0, su1)
        assert float(rmax1.item()) == rmax0 and rmax0 > 0.0
