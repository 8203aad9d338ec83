"""Device tensors handed to the solver straight from torch kernels.  Needs an
MI355X.

The library runs on its own non-blocking stream; the wrapper drains torch's
current stream before a device tensor crosses the C-ABI.  A right-hand side
computed by a long queue of torch kernels must therefore give exactly the
solve that the same values uploaded from the host give.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_gpu_rhs_from_pending_torch_kernels_matches_host_upload(gpu_lib):
    import torch
    from hydro3d_amd import MultigridSolver
    n = 64
    spec = dict(n=(n,) * 3, h=(1.0 / n,) * 3, bc=(1,) * 6, nlev=4)
    rng = np.random.default_rng(5)
    f0 = rng.standard_normal((n, n, n))
    g = torch.from_numpy(f0).cuda()
    torch.cuda.synchronize()
    # queue enough work that the last kernel is still pending at set_rhs
    big = torch.ones(4096, 4096, dtype=torch.float64, device="cuda")
    for _ in range(4):
        big = big @ big * (1.0 / 4096)
    print("case device rhs 64^3 periodic", flush=True)
    with MultigridSolver([spec], tol=0.0, maxcy=3) as S:
        f = (g + 0.0 * big[0, 0]) - g.mean()
        S.set_rhs(0, f)
        S.solve()
        p_dev = S.get_p(0)
    ref = f.cpu().numpy()  # the values the device tensor ends up holding
    with MultigridSolver([spec], tol=0.0, maxcy=3) as S:
        S.set_rhs(0, ref)
        S.solve()
        p_host = S.get_p(0)
    assert np.array_equal(p_dev, p_host)


def test_gpu_p_into_pending_torch_memory(gpu_lib):
    """get_p into a tensor torch is still writing: the library's copy lands
    after torch's kernels, so the tensor holds the pressure."""
    import torch
    from hydro3d_amd import MultigridSolver
    n = 32
    spec = dict(n=(n,) * 3, h=(1.0 / n,) * 3, bc=(1,) * 6, nlev=3)
    f = np.random.default_rng(6).standard_normal((n, n, n))
    f -= f.mean()
    with MultigridSolver([spec], tol=0.0, maxcy=2) as S:
        S.set_rhs(0, f)
        S.solve()
        p_host = S.get_p(0)[1:-1, 1:-1, 1:-1]
        out = torch.empty((n, n, n), dtype=torch.float64, device="cuda")
        big = torch.ones(4096, 4096, dtype=torch.float64, device="cuda")
        for _ in range(4):
            big = big @ big * (1.0 / 4096)
        out.fill_(1.0).add_(big[0, 0])  # pending writes to `out`
        S.get_p(0, out=out)
        assert np.array_equal(out.cpu().numpy(), p_host)
