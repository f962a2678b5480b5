"""Critical points in parts on the device (tg_critpoints_split): the
reference's partition loop with device localisation
(src/alg/critPoints.cu:654-1003, merged at :1438-1491), and the automatic
split of a mesh whose footprint exceeds the device budget, which must give
the whole mesh's classes."""
import numpy as np
import pytest

import tetragpu_amd as t
from oracle import binding as O

pytestmark = pytest.mark.gpu


def resident(T, V, f64=True):
    """The split's resident arrays (TV, scalars, classes, cell and vertex
    positions, ring marks) with a little slack for its scan scratch."""
    return 16 * T + (8 if f64 else 4) * V + 12 * V + 4 * (T + 1) + 4 * (V + 1) + V + (18 << 10)


@pytest.mark.parametrize("spec", [
    dict(n=24, seed=3, permute_vertices=False, permute_cells=False, scalar="gaussians"),
    dict(n=20, seed=4, permute_vertices=False, permute_cells=True, split_fraction=0.2,
         n_hubs=2, hub_splits=30, scalar="sines"),
    dict(n=20, seed=5, scalar="uniform"),   # permuted: wide halos, many small parts
], ids=["ordered", "splits_hubs", "permuted"])
def test_auto_split_equals_whole_mesh(gpu, spec):
    m = t.kuhn(**spec)
    T, V = m.n_cells, m.n_points
    s = m.scalars.astype(np.float64)
    whole = O.critpoints(O.vv(m.cells, V), s, V)
    full = t.vv_footprint(T, V)
    print("split", spec, "T", T, "footprint", full, flush=True)
    for div in (2, 5):
        budget = resident(T, V) + full // div
        got, st = t.split_classes(m.cells, V, s, budget_bytes=budget)
        print("  div", div, st, flush=True)
        assert st["parts"] >= 2 and st["budget_bytes"] == budget
        assert np.array_equal(got, whole)


def test_whole_mesh_when_it_fits(gpu):
    m = t.kuhn(12, seed=9, scalar="gaussians")
    got, st = t.split_classes(m.cells, m.n_points, m.scalars)   # default budget: free HBM
    assert st["parts"] == 1 and st["cells_total"] == m.n_cells
    assert np.array_equal(got, O.critpoints(O.vv(m.cells, m.n_points),
                                            m.scalars.astype(np.float64), m.n_points))


def test_float32_scalars_and_isolated_vertex(gpu):
    m = t.kuhn(16, seed=2, permute_vertices=False, permute_cells=False, scalar="gaussians")
    V = m.n_points + 2   # two vertices in no cell
    s = np.append(m.scalars, [0.25, 0.75]).astype(np.float32)
    whole = O.critpoints(O.vv(m.cells, V), s, V)
    assert whole[-2:].tolist() == [[0, 0, 4], [0, 0, 4]]
    got, st = t.split_classes(m.cells, V, s, budget_bytes=resident(m.n_cells, V, False) +
                              t.vv_footprint(m.n_cells, V) // 4)
    assert st["parts"] >= 2 and np.array_equal(got, whole)


@pytest.mark.parametrize("with_rows", [False, True])
def test_caller_partitions_equal_oracle(gpu, with_rows):
    m = t.kuhn(14, seed=6, split_fraction=0.15, scalar="gaussians")
    rng = np.random.default_rng(3)
    V, T = m.n_points, m.n_cells
    parts = rng.integers(0, 5, V).astype(np.int32)
    rows = None
    if with_rows:
        # the vertex rule's rows with a few cells dropped and a few added
        rows = np.zeros((5, T), np.int32)
        for p in range(5):
            rows[p] = (parts[m.cells] == p).any(axis=1)
        flip = rng.random(rows.shape) < 0.02
        rows = np.where(flip, 1 - rows, rows).astype(np.int32)
    s = m.scalars.astype(np.float64)
    got, st = t.split_classes(m.cells, V, s, parts=parts, part_cells=rows)
    assert st["parts"] == 5
    assert np.array_equal(got, O.critpoints_partitioned(m.cells, V, s, parts, rows))


def test_partition_too_large_for_budget(gpu):
    m = t.kuhn(20, seed=1)
    with pytest.raises(t.mesh._native.NativeError) as e:
        t.split_classes(m.cells, m.n_points, m.scalars, parts=np.zeros(m.n_points, np.int32),
                        budget_bytes=resident(m.n_cells, m.n_points) + 4096)
    assert e.value.status == -3 and "does not fit" in str(e.value)
    with pytest.raises(t.mesh._native.NativeError) as e:
        t.split_classes(m.cells, m.n_points, m.scalars, budget_bytes=4096)
    assert e.value.status == -3 and "exceed the device budget" in str(e.value)


def test_partitioned_classes_runs_on_device(gpu):
    # the host library's partition loop entry (tetra_partitioned_classes)
    # goes through the device localisation
    m = t.kuhn(12, seed=8, scalar="gaussians")
    parts = (np.arange(m.n_points) * 7 % 4).astype(np.int32)
    s = m.scalars.astype(np.float64)
    got, st = t.partitioned_classes(m.cells, m.n_points, s, parts)
    assert st["partitions"] == 4
    assert np.array_equal(got, O.critpoints_partitioned(m.cells, m.n_points, s, parts))


def test_ids_outside_the_mesh_rejected(gpu):
    m = t.kuhn(8, seed=4, permute_vertices=False, permute_cells=False)
    bad = m.cells.copy()
    bad[100, 2] = m.n_points + 5
    for kw in (dict(budget_bytes=resident(m.n_cells, m.n_points) +
                    t.vv_footprint(m.n_cells, m.n_points) // 3),
               dict(parts=np.zeros(m.n_points, np.int32))):
        with pytest.raises(t.mesh._native.NativeError) as e:
            t.split_classes(bad, m.n_points, m.scalars, **kw)
        assert e.value.status == -4 and "cell 100" in str(e.value)
