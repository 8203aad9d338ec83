"""The reference build on several MPI ranks (oracle/_ref/ref_multiblock,
mode timeloop): the CPU baseline's multi-rank runs.  A 2x2x2 box of blocks
dealt out over 1, 2, 4 and 8 ranks gives the same final rmax -- the
reference's exchange between ranks moves the same values as between the
blocks of one rank.  No GPU; skipped where the reference build or
mpiexec is missing."""
import os

import pytest

MPIEXEC = "/opt/conda/bin/mpiexec"


def test_reference_loop_independent_of_rank_count():
    from hydro3d_amd import configs as C
    from oracle import oracle as O
    if not O.multiblock_available() or not os.path.exists(MPIEXEC):
        pytest.skip("reference multi-block driver or mpiexec not available")
    c = C.cube(16, C.BC_PPN, 3, seed=0)
    gdx = c["h"]
    doms = O.box_layout([8, 8], [8, 8], [8, 8], gdx)
    ins = O.block_inputs(doms, c["u"], c["v"], c["w"], gdx)
    ncores = len(os.sched_getaffinity(0))
    rmax = {}
    for nr in (1, 2, 4, 8):
        if nr > ncores:
            continue
        _, rmax[nr] = O.time_multiblock(doms, ins, gdx, c["bc"], 2, 3, nr, timeout=120)
    assert len(set(rmax.values())) == 1, rmax
