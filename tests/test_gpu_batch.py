"""Batches of independent meshes (tg_batch_*): every mesh's VV rows and
classes equal the oracle's for that mesh alone, across wave boundaries, both
kernel forms, and hipGraph replays (each replay checked against the eager
result, with eager work and synchronisation between replays)."""
import numpy as np
import pytest

from oracle import binding as O

pytestmark = pytest.mark.gpu


def mixed_meshes(seed, count):
    import tetragpu_amd as t
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        kind = i % 4
        if kind == 3:  # a non-manifold soup
            V = int(rng.integers(8, 60))
            T = int(rng.integers(1, 80))
            cells = np.array([rng.choice(V, 4, replace=False) for _ in range(T)], np.int32)
            out.append((cells, V, rng.integers(0, 5, V).astype(np.float32)))
        else:
            n = int(rng.integers(1, 7))
            m = t.kuhn(n, seed=int(rng.integers(1 << 30)), order=["sorted", "path", "shuffled"][kind],
                       scalar=["uniform", "gaussians", "sines"][kind])
            out.append((m.cells, m.n_points, m.scalars))
    out.append((np.zeros((0, 4), np.int32), 3, np.zeros(3, np.float32)))  # empty mesh
    return out


def check_batch(b, meshes):
    cls = b.classes()
    for i, (cells, V, s) in enumerate(meshes):
        VV = O.vv(cells, V)
        ref = O.critpoints(VV, np.asarray(s, np.float64), V)
        assert np.array_equal(b.mesh_classes(i, cls), ref), f"mesh {i} classes"
        off, vals = b.vv(i)
        srt = VV.sorted_rows()
        assert np.array_equal(off, srt.off) and np.array_equal(vals, srt.val[:, 0]), f"mesh {i} VV"


@pytest.mark.parametrize("variant", [0, 1, 2, 3, 4])
def test_batch_matches_per_mesh_oracle(gpu, variant):
    meshes = mixed_meshes(1, 40)
    print(f"batch variant={variant} meshes={len(meshes)}", flush=True)
    with gpu.Batch(meshes) as b:
        if variant:
            b.set_config(variant=variant)
        b.vv_critpoints()
        b.sync()
        check_batch(b, meshes)


def test_batch_many_waves(gpu):
    meshes = mixed_meshes(2, 30)
    with gpu.Batch(meshes, max_wave_cells=2000) as b:
        assert b.counts()["waves"] > 3
        b.vv_critpoints()
        check_batch(b, meshes)


def test_batch_graph_replays(gpu):
    meshes = mixed_meshes(3, 24)
    with gpu.Batch(meshes, max_wave_cells=3000) as b:
        b.vv_critpoints()
        eager = b.classes()
        eager_vv = [b.vv(i) for i in range(len(meshes))]
        b.use_graph(True)
        for rep in range(4):
            print(f"graph replay {rep}", flush=True)
            b.vv_critpoints()          # rep 0: eager warm-up, 1: capture, 2+: replay
            b.sync()
            assert np.array_equal(b.classes(), eager), f"replay {rep} classes"
            for i in (0, len(meshes) // 2, len(meshes) - 2):
                off, vals = b.vv(i)
                assert np.array_equal(off, eager_vv[i][0]) and np.array_equal(vals, eager_vv[i][1])
        check_batch(b, meshes)


def test_batch_rejects_cross_mesh_ids(gpu):
    from tetragpu_amd._native import NativeError
    good = (np.array([[0, 1, 2, 3]], np.int32), 4, np.zeros(4, np.float32))
    bad = (np.array([[0, 1, 2, 5]], np.int32), 4, np.zeros(4, np.float32))  # 5 is not in mesh 1
    with pytest.raises(NativeError) as e:
        gpu.Batch([good, bad])
    assert e.value.status == -4
