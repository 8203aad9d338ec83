"""The drop-in itself: `newsolv_mg` as the LES driver calls it
(flosol.for:215), bound to hydro3d_amd/fortran/newsolv_mg.f90 ->
newsolv_mg_h3d.f90 -> the device, run on the reference's own module state
(oracle/_ref/dropin_driver, set up by mb_setup.f90 exactly as the reference
loop's driver ref_multiblock sets it up: read_mdmap, read_infodom,
localparameters' sizes, coeff).  Needs an MI355X and the reference build
(oracle/_ref, built where the reference tree is present).

Each case runs the reference's newsolv_mg loop (newsolv_mg.for:17-59, with
its exit at rmax < eps .and. iter > 1) and the drop-in from the same u*,
v*, w* and p = 0, and compares bitwise: iter, rmax, every block's p with its
ghost layer, su, and the whole u, v, w arrays (ghost layers included, after
the drop-in's exchange(1..4) and pbound on the host).
"""
import numpy as np
import pytest

from h3d_util import load_golden

pytestmark = pytest.mark.gpu


def _need_drivers():
    from oracle import oracle as O
    if not (O.multiblock_available() and O.dropin_available()):
        pytest.skip("reference-linked drivers (oracle/_ref) not built")
    return O


def _inputs(doms, bcs, seed, gdx, cds4=False):
    """Per block u*, v*, w* (faces 0..n of their axis, or -1..n+1 with cds4:
    the domain's face arrays extended by configs.extend_faces) and p = 0."""
    from hydro3d_amd import configs as C
    N = [0, 0, 0]
    for d in doms:
        for a in range(3):
            N[a] = max(N[a], int(round(d["hi"][a] / gdx[a])))
    u, v, w = C.face_velocity(tuple(N), bcs, seed, 0.0, gdx)
    e = 0
    if cds4:
        u, v, w = C.extend_faces(u, v, w, bcs, seed + 1)
        e = 1
    ins = []
    for d in doms:
        x0, y0, z0 = (int(round(d["lo"][a] / gdx[a])) for a in range(3))
        ni, nj, nk = d["n"]
        ins += [u[z0:z0 + nk, y0:y0 + nj, x0:x0 + ni + 1 + 2 * e],
                v[z0:z0 + nk, y0:y0 + nj + 1 + 2 * e, x0:x0 + ni],
                w[z0:z0 + nk + 1 + 2 * e, y0:y0 + nj, x0:x0 + ni], np.zeros((nk, nj, ni))]
    return ins


def _split(out, doms, pl):
    """iter, rmax, per block (p padded, su), per block (u, v, w) arrays."""
    it, rmax = int(out[0]), float(out[1])
    q = 2
    blocks = []
    for d in doms:
        ni, nj, nk = d["n"]
        npd, nin = (ni + 2) * (nj + 2) * (nk + 2), ni * nj * nk
        p = out[q:q + npd].reshape(nk + 2, nj + 2, ni + 2)
        q += npd
        su = out[q:q + nin].reshape(nk, nj, ni)
        q += nin
        blocks.append([p, su])
    for b, d in enumerate(doms):
        ni, nj, nk = d["n"]
        nt = (ni + 2 * pl) * (nj + 2 * pl) * (nk + 2 * pl)
        for _ in range(3):
            blocks[b].append(out[q:q + nt])
            q += nt
    assert q == out.size, (q, out.size)
    return it, rmax, blocks


def _compare(O, doms, gdx, bcs, nlev, maxcy, eps, seed, nu2=2, cds4=False):
    from oracle.oracle import infodom_text, mdmap_text
    info, mdmap = infodom_text(doms), mdmap_text(len(doms))
    ins = _inputs(doms, bcs, seed, gdx, cds4)
    kw = dict(nlev=nlev, ncycles=maxcy, eps=eps, dt=1e-3, nu2=nu2, early=1, timeout=900, cds4=cds4)
    ref = O.run_multiblock_files(info, mdmap, "newsolv", ins, None, gdx, bcs, **kw)
    print("reference loop done", flush=True)
    dev = O.run_multiblock_files(info, mdmap, "newsolv", ins, None, gdx, bcs,
                                 driver=O.DROPIN_DRIVER, **kw)
    print("drop-in done", flush=True)
    it_r, rm_r, br = _split(ref, doms, 2)
    it_d, rm_d, bd = _split(dev, doms, 2)
    print(f"iter {it_r} / {it_d}, rmax {rm_r:.6e} / {rm_d:.6e}", flush=True)
    assert it_d == it_r and rm_d == rm_r
    names = ("p", "su", "u", "v", "w")
    for b in range(len(doms)):
        for q, nm in enumerate(names):
            assert np.array_equal(bd[b][q], br[b][q]), f"block {b} {nm}"


@pytest.mark.parametrize("name", ["mb_kh.npz", "mb_box.npz", "mb_perx.npz"])
def test_gpu_dropin_newsolv_mg_bitwise_vs_reference_loop(gpu_lib, name):
    O = _need_drivers()
    z = load_golden(name)
    doms = [dict(id=d, n=tuple(int(x) for x in z["n"][d]), lo=tuple(float(x) for x in z["lo"][d]),
                 hi=tuple(float(x) for x in z["hi"][d])) for d in range(len(z["n"]))]
    gdx = tuple(float(x) for x in z["h"])
    bcs = tuple(int(x) for x in z["bc"])
    print(f"case drop-in {name}: {len(doms)} blocks", flush=True)
    _compare(O, doms, gdx, bcs, int(z["nlev"]), 30, 1e-6, seed=901)


@pytest.mark.parametrize("name", ["mb_kh.npz", "mb_box.npz", "mb_perx.npz"])
def test_gpu_dropin_cds4_bitwise_vs_reference_loop(gpu_lib, name):
    """differencing = 2 (calmas' 4th-order divergence, press.for:17-37, the
    setting of four shipped benchmarks): the drop-in against the reference's
    loop on the same layouts, bitwise as above; the stencil's second ghost
    faces come from the neighbour or partner across a shared face and from
    the host arrays at walls."""
    O = _need_drivers()
    z = load_golden(name)
    doms = [dict(id=d, n=tuple(int(x) for x in z["n"][d]), lo=tuple(float(x) for x in z["lo"][d]),
                 hi=tuple(float(x) for x in z["hi"][d])) for d in range(len(z["n"]))]
    gdx = tuple(float(x) for x in z["h"])
    bcs = tuple(int(x) for x in z["bc"])
    print(f"case drop-in cds4 {name}: {len(doms)} blocks", flush=True)
    _compare(O, doms, gdx, bcs, int(z["nlev"]), 30, 1e-6, seed=911, cds4=True)


def test_gpu_dropin_single_block(gpu_lib):
    """One 32^3 block, periodic x/y onto itself, walls in z (the 64^3
    workload's boundaries at 32^3): the shim's single-block path."""
    O = _need_drivers()
    gdx = (1.0 / 32,) * 3
    doms = O.box_layout([32], [32], [32], gdx)
    print("case drop-in single 32^3 block", flush=True)
    _compare(O, doms, gdx, (1, 1, 1, 1, 0, 0), 3, 30, 1e-6, seed=902)


def test_gpu_dropin_kelvin_helmholtz_full_size(gpu_lib):
    """The Kelvin-Helmholtz benchmark's own layout (infodom.cin: 8 x 1 x 8
    blocks of 64 x 8 x 64 cells, periodic x with partners, periodic y onto
    itself, walls in z) and solver settings (control.cin: 4 levels, V(2,1),
    red-black, at most 30 iterations, eps 1e-6), one rank, from a seeded
    velocity field."""
    O = _need_drivers()
    from oracle.oracle import box_layout
    gdx = (0.0025,) * 3
    doms = box_layout([64] * 8, [8], [64] * 8, gdx)
    print("case drop-in KH 64 blocks of 64x8x64", flush=True)
    _compare(O, doms, gdx, (1, 1, 1, 1, 0, 0), 4, 30, 1e-6, seed=903, nu2=1)
