"""The comparison with a TTK / ParaView export (tetragpu_amd/ttk.py): the
export format of convert_vtk_to_txt.py:8-59 (integer or named classes,
regular points not logged) read back exactly, the six-way and four-way
agreement, and the reader's errors.  TTK's own algorithm is not available
here: agreement with TTK itself stays unpinned."""
import numpy as np
import pytest

import tetragpu_amd as t
from oracle import binding as O
from tetragpu_amd import ttk


@pytest.fixture(scope="module")
def labelled():
    m = t.kuhn(6, seed=21, split_fraction=0.2, scalar="gaussians")
    cls = O.critpoints(O.vv(m.cells, m.n_points), m.scalars.astype(np.float64), m.n_points)
    return cls, O.label6(cls)


@pytest.mark.parametrize("as_string", [False, True])
def test_round_trip_agrees(tmp_path, labelled, as_string):
    cls, lab = labelled
    p = tmp_path / "ttk.txt"
    n = ttk.write_export(p, lab, as_string=as_string)
    assert n == int((lab != 5).sum()) > 0
    types = ttk.read_export(p, len(lab))
    assert np.array_equal(types, lab)
    r = ttk.compare(lab, types, cls)
    assert r["mismatch6"] == 0 and r["mismatch4"] == 0 and r["agree6"] == len(lab)
    assert np.trace(np.array(r["confusion6"])) == len(lab)


def test_pyclassify_grouping(labelled):
    # pyclassify.py:65 groups TTK's types as [1,4,4,2,4,3]; on our own labels
    # that grouping must give back the numeric class of every point
    cls, lab = labelled
    assert np.array_equal(ttk.CLASS4_OF_TTK[lab], cls[:, 2].astype(np.uint8))


def test_disagreements_are_reported(labelled):
    cls, lab = labelled
    ttk_types = lab.copy()
    sad = np.flatnonzero(lab == 1)[:2]
    mins = np.flatnonzero(lab == 0)[:1]
    ttk_types[sad] = 2          # 1-saddle vs 2-saddle: six-way only
    ttk_types[mins] = 5         # a minimum TTK calls regular: both levels
    r = ttk.compare(lab, ttk_types, cls)
    assert r["mismatch6"] == len(sad) + len(mins)
    assert r["mismatch4"] == len(mins)
    assert r["first_mismatch6"] == sorted(np.concatenate([sad, mins]).tolist())
    assert r["confusion6"][1][2] == len(sad) and r["confusion6"][0][5] == len(mins)


def test_reader_errors(tmp_path):
    p = tmp_path / "bad.txt"
    p.write_text("3 1\n4 1 7\n")
    with pytest.raises(ValueError, match="expected '<index> <class>'"):
        ttk.read_export(p, 10)
    p.write_text("3 9\n")
    with pytest.raises(ValueError, match="not a TTK critical type"):
        ttk.read_export(p, 10)
    p.write_text("12 Maxima\n")
    with pytest.raises(ValueError, match="outside"):
        ttk.read_export(p, 10)
    p.write_text("2 Maxima\n5 Multi-Saddle\n\n")
    assert ttk.read_export(p, 6).tolist() == [5, 5, 3, 5, 5, 4]
