"""The multi-block layout reader (h3dmg_layout_*, hydro3d_amd/layout.py)
against the reference's own reading of the same files: read_infodom's
neighbour and periodic-partner search (alloc_dom.for:524-690,1040-1215) as
the reference build ran it on the shipped benchmarks' infodom.cin and on the
fixture layouts (tests/golden/mb_*.npz).  Host only."""
import os

import numpy as np
import pytest

from h3d_util import GOLDEN, load_golden
from hydro3d_amd import _lib as L
from hydro3d_amd.layout import Layout

BENCH = load_golden("mb_benchmark_layouts.npz")
NAMES = [str(x) for x in BENCH["names"]]


def _write(tmp_path, name, text):
    p = tmp_path / name
    p.write_text(str(text))
    return str(p)


def _one_rank(ndoms):
    return (" 1  (1)sub-domains description\n"
            f"  {ndoms} number of domain\n  1 number of processor\n ====\n"
            f"  0  {ndoms}  " + " ".join(str(i) for i in range(ndoms)) + "\n")


@pytest.mark.parametrize("name", NAMES)
def test_benchmark_neighbours_equal_the_reference_search(tmp_path, name):
    info = _write(tmp_path, "infodom.cin", BENCH[f"{name}_infodom"])
    mdmap = _write(tmp_path, "mdmap.cin", BENCH[f"{name}_mdmap"])
    codes = [int(x) for x in BENCH[f"{name}_codes"]]
    tab = BENCH[f"{name}_nbr_table"]
    with Layout(info, mdmap) as lay:
        assert lay.ndomains == tab.shape[0]
        for d in range(lay.ndomains):
            pn, pp = lay.neighbours(d, codes)
            row = tab[d]
            assert row[0] == d
            assert list(pn) == list(row[4:10]), (name, d, pn, row[4:10])
            assert list(pp) == list(row[30:36]), (name, d, pp, row[30:36])


@pytest.mark.parametrize("name", NAMES)
def test_benchmark_blocks_per_rank(tmp_path, name):
    """mdmap.cin's type-1 rank lists: every domain on exactly one rank, the
    cells from the coordinates and control.cin's spacing as the reference
    derived them (the neighbour table's cell counts)."""
    info = _write(tmp_path, "infodom.cin", BENCH[f"{name}_infodom"])
    mdmap = _write(tmp_path, "mdmap.cin", BENCH[f"{name}_mdmap"])
    gdx = [float(x) for x in BENCH[f"{name}_gdx"]]
    codes = [int(x) for x in BENCH[f"{name}_codes"]]
    tab = BENCH[f"{name}_nbr_table"]
    seen = []
    with Layout(info, mdmap) as lay:
        for r in range(lay.nranks):
            try:
                specs, ids = lay.blocks(r, gdx, codes, 1)
            except L.H3dError as e:
                # a neighbour on another rank: the multi-rank exchange is
                # not on the device path
                assert e.status == L.ERR_UNSUPPORTED and "rank" in str(e)
                continue
            for s, d in zip(specs, ids):
                assert tuple(s["n"]) == tuple(int(x) for x in tab[d][1:4])
            seen += ids
        # all domains on one rank: every block is placed, with its neighbours
        one = _write(tmp_path, "one.cin", _one_rank(lay.ndomains))
    with Layout(info, one) as lay:
        specs, ids = lay.blocks(0, gdx, codes, 1)
        assert ids == list(range(lay.ndomains))
        for d, s in enumerate(specs):
            pn = tab[d][4:10]
            for f in range(6):
                if pn[f] >= 0:
                    assert s["bc"][f] == L.BC_NEIGHBOUR and s["nbr"][f] == pn[f]


def test_kelvin_helmholtz_layout():
    """The KH benchmark: 8 x 1 x 8 blocks of 64 x 8 x 64 cells, periodic
    along x (partners at the ends) and y (one block spans it), walls in z;
    rank 0 holds every eighth domain (mdmap.cin)."""
    info = str(BENCH["Kelvin_Helmholtz_infodom"])
    mdmap = str(BENCH["Kelvin_Helmholtz_mdmap"])
    import tempfile
    with tempfile.TemporaryDirectory() as td:
        pi, pm = os.path.join(td, "i.cin"), os.path.join(td, "m.cin")
        open(pi, "w").write(info)
        open(pm, "w").write(mdmap)
        with Layout(pi, pm) as lay:
            assert lay.ndomains == 64 and lay.nranks == 8
            assert [lay.domain(d)[0] for d in (0, 8, 56, 1, 63)] == [0, 0, 0, 1, 7]
            one = os.path.join(td, "one.cin")
            open(one, "w").write(_one_rank(64))
        with Layout(pi, one) as lay:
            specs, _ = lay.blocks(0, (0.0025,) * 3, (5, 5, 5, 5, 3, 3), 4)
    assert all(s["n"] == (64, 8, 64) for s in specs)
    s0, s7 = specs[0], specs[7]
    assert s0["bc"] == (2, 2, 1, 1, 0, 2) and s0["nbr"][0] == 7 and s0["nbr"][1] == 1
    assert s7["nbr"][1] == 0 and s0["nbr"][5] == 8


def test_fixture_layouts_round_trip(tmp_path):
    """The fixture layouts written by the oracle's infodom writer read back
    to the reference's neighbour table."""
    from oracle import oracle as O
    for path in sorted(os.listdir(GOLDEN)):
        if not path.startswith("mb_") or "benchmark" in path:
            continue
        z = load_golden(path)
        doms = [dict(id=d, n=tuple(int(x) for x in z["n"][d]), lo=tuple(z["lo"][d]),
                     hi=tuple(z["hi"][d])) for d in range(len(z["n"]))]
        info = _write(tmp_path, "infodom.cin", O.infodom_text(doms))
        mdmap = _write(tmp_path, "mdmap.cin", O.mdmap_text(len(doms)))
        codes = [5 if int(b) == 1 else 1 for b in z["bc"]]
        with Layout(info, mdmap) as lay:
            for d in range(lay.ndomains):
                pn, pp = lay.neighbours(d, codes)
                assert list(pn) == list(z["nbr_table"][d][4:10])
                assert list(pp) == list(z["nbr_table"][d][30:36])


def test_bad_files_are_refused(tmp_path):
    with pytest.raises(L.H3dError) as e:
        Layout(str(tmp_path / "missing.cin"), str(tmp_path / "missing2.cin"))
    assert e.value.status == L.ERR_ARG
    info = _write(tmp_path, "infodom.cin", BENCH["CavityFlow_infodom"])
    bad = _write(tmp_path, "mdmap.cin", " 1\n  3 number of domain\n 1\n ===\n 0 3 0 1 2\n")
    with pytest.raises(L.H3dError) as e:
        Layout(info, bad)
    assert "infodom.cin has 8" in str(e.value)
