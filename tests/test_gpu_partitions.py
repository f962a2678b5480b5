"""The partition loop on the device (src/alg/critPoints.cu:654-1003, merged
as at :1438-1491) through the package's partitioned_classes, against the
oracle's restatement (oracle.binding.critpoints_partitioned) and, on Kuhn
grids (flag complexes), against the whole-mesh classes."""
import numpy as np
import pytest

import tetragpu_amd as t
from oracle import binding as O

pytestmark = pytest.mark.gpu


def slabs(m, k):
    z = m.positions[:, 2].astype(np.float64)
    return np.minimum(((z - z.min()) / (z.max() - z.min()) * k).astype(np.int32), k - 1)


@pytest.mark.parametrize("k", [1, 3, 8])
def test_slabs_equal_oracle(gpu, k):
    m = t.kuhn(12, seed=k, split_fraction=0.1, scalar="gaussians")
    parts = slabs(m, k)
    sc = m.scalars.astype(np.float64)
    print(f"slabs k={k} V={m.n_points} T={m.n_cells}", flush=True)
    cls, st = t.partitioned_classes(m.cells, m.n_points, sc, parts)
    assert st["partitions"] == k
    assert np.array_equal(cls, O.critpoints_partitioned(m.cells, m.n_points, sc, parts))


def test_random_partitions_of_soup_equal_oracle(gpu):
    rng = np.random.default_rng(12)
    V, T = 300, 900
    cells = np.array([rng.choice(V, 4, replace=False) for _ in range(T)])
    sc = rng.integers(0, 9, V).astype(np.float64)
    parts = rng.integers(0, 7, V)
    cls, _ = t.partitioned_classes(cells, V, sc, parts)
    assert np.array_equal(cls, O.critpoints_partitioned(cells, V, sc, parts))


def test_kuhn_slabs_equal_whole_mesh(gpu):
    m = t.kuhn(40, seed=3, scalar="gaussians")
    with t.Mesh(m.cells, m.n_points, m.scalars.astype(np.float64)) as d:
        d.vv_critpoints()
        whole = d.classes()
    cls, st = t.partitioned_classes(m.cells, m.n_points, m.scalars.astype(np.float64),
                                    slabs(m, 6))
    print(f"kuhn40 slabs stats {st}", flush=True)
    assert np.array_equal(cls, whole)
    assert st["max_cells"] < m.n_cells


def test_bad_arguments_rejected(gpu):
    with pytest.raises(ValueError):
        t.partitioned_classes(np.zeros((1, 4), np.int32), 4, np.zeros(4), np.array([0, 0, -1, 0]))
    with pytest.raises(ValueError):
        t.partitioned_classes(np.zeros((1, 4), np.int32), 4, np.zeros(3), np.zeros(4))


def test_invalid_cells_raise(gpu):
    from tetragpu_amd._native import NativeError

    with pytest.raises(NativeError) as e:
        t.partitioned_classes(np.array([[0, 1, 1, 2]], np.int32), 4, np.zeros(4), np.zeros(4))
    assert e.value.status == -4
    with pytest.raises(ValueError, match="outside"):
        t.partitioned_classes(np.array([[0, 1, 2, 7]], np.int32), 4, np.zeros(4), np.zeros(4))
