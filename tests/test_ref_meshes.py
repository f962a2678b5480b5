"""The reference's own meshes (mwe, tiny, big of artificial_dataset.sh:6-28,
built by make_mesh.py's fan rule) as oracle fixtures.

- Regeneration equals the committed fixture (cells and scalars of mwe and
  tiny in full, SHA-256 digests for big).
- Closed forms of a fan mesh: every cell after the first adds one point,
  three edges and three faces, so T = V - 3, E = 3T + 3, F = 3T + 1 and
  V - E + F - T = 1; no face lies in more than two cells.
- mwe classified by hand: cells (0,1,2,3), (0,1,2,4); scalars
  RandomState(1234).rand(5) = .192 .622 .438 .785 .780.  Vertex 0 is below
  all four neighbours (a minimum); 3 and 4 are above theirs (maxima); 1 has
  lower {0, 2} (one component: 0-2 is an edge) and upper {3, 4} (two: 3-4 is
  not an edge) -- a saddle; 2 has lower {0} and upper {1, 3, 4} joined
  through 1 -- regular.
"""
import json
import sys
from pathlib import Path

import numpy as np
import pytest

from oracle import binding as O

sys.path.insert(0, str(Path(__file__).resolve().parent / "golden"))
import ref_meshes as R  # noqa: E402

FIX = json.loads(R.FIXTURE.read_text())


def load_mesh(name):
    """(cells, n_points, scalars float64) from the committed fixture files."""
    return R.load(name)


@pytest.mark.parametrize("name", ["mwe", "tiny", "big"])
def test_fixture_files(name):
    """The committed cells and scalars match the fixture's digests."""
    cells, V, sc = load_mesh(name)
    f = FIX[name]
    assert V == f["n_points"] and cells.shape[0] == f["n_cells"]
    assert R.digest(cells.astype("<i4")) == f["cells_sha256"]
    assert R.digest(sc.astype("<f8")) == f["scalars_sha256"]
    if "cells" in f:
        assert [float.fromhex(x) for x in f["scalars"]] == sc.tolist()
    else:
        for i, row in f["sample_cells"].items():
            assert cells[int(i)].tolist() == row


@pytest.mark.parametrize("name,prefix", [("mwe", None), ("tiny", None), ("big", 4000)])
def test_regenerated_equals_fixture(name, prefix):
    """The restated rule reproduces the committed cells (all of mwe and tiny,
    the first 4000 cells of big: the whole of big takes minutes)."""
    n = R.MESHES[name]
    got = R.fan_cells(n, prefix)
    want = load_mesh(name)[0]
    assert np.array_equal(got, want[: got.shape[0]])


@pytest.mark.parametrize("name", ["mwe", "tiny", "big"])
def test_fan_closed_forms(name):
    cells, V, _ = load_mesh(name)
    T = cells.shape[0]
    E, _, VE = O.edges(cells, V)
    F, TF, _ = O.faces(cells, V)
    assert T == V - 3 and E == 3 * T + 3 and F == 3 * T + 1
    assert V - E + F - T == 1
    O.ft(TF, F)  # raises when a face lies in more than two cells
    assert (np.diff(cells, axis=1) > 0).all()  # cells ascending, as the rule builds them


def test_mwe_by_hand():
    cells, V, sc = load_mesh("mwe")
    assert cells.tolist() == [[0, 1, 2, 3], [0, 1, 2, 4]]
    VV = O.vv(cells, V).sorted_rows()
    rows = [sorted(VV.row(v)[:, 0].tolist()) for v in range(V)]
    assert rows == [[1, 2, 3, 4], [0, 2, 3, 4], [0, 1, 3, 4], [0, 1, 2], [0, 1, 2]]
    cls = O.critpoints(VV, sc, V)
    assert cls.tolist() == [[0, 1, 1], [1, 2, 4], [1, 1, 3], [1, 0, 2], [1, 0, 2]]
