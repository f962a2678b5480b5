"""Synthetic Kuhn meshes: closed-form counts, determinism, permutations."""
import numpy as np
import pytest

import tetragpu_amd as t


@pytest.mark.parametrize("n", [1, 2, 5])
def test_kuhn_counts(n):
    m = t.kuhn(n)
    c = t.kuhn_counts(n)
    assert m.n_points == c["V"] and m.n_cells == c["T"]
    assert m.cells.min() >= 0 and m.cells.max() < m.n_points
    # ascending corners, all distinct
    assert (np.diff(m.cells, axis=1) > 0).all()
    assert m.max_vt == (24 if n >= 2 else 6)


def test_seeded_and_deterministic():
    a, b, c = t.kuhn(6, seed=1), t.kuhn(6, seed=1), t.kuhn(6, seed=2)
    assert np.array_equal(a.cells, b.cells) and np.array_equal(a.scalars, b.scalars)
    assert not np.array_equal(a.cells, c.cells)


def test_uniform_scalars_exact_f32():
    m = t.kuhn(4)
    k = m.scalars.astype(np.float64) * 2 ** 24
    assert np.array_equal(k, np.round(k)) and (m.scalars < 1).all() and (m.scalars >= 0).all()


def test_unpermuted_linear_field_is_grid_index():
    m = t.kuhn(3, permute_vertices=False, permute_cells=False, jitter=False, scalar="linear")
    assert np.array_equal(m.scalars, np.arange(64, dtype=np.float32))


def test_splits_and_hubs():
    base = t.kuhn(6)
    m = t.kuhn(6, split_fraction=0.15, n_hubs=2, hub_splits=50)
    n_split = round(0.15 * base.n_cells) + 2 * 50
    assert m.n_cells == base.n_cells + 3 * n_split
    assert m.n_points == base.n_points + n_split
    assert m.max_vt > 100


def test_bad_arguments():
    with pytest.raises(ValueError):
        t.kuhn(0)
    with pytest.raises(ValueError):
        t.kuhn(4, split_fraction=0.1, scalar="linear")
