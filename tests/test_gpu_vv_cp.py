"""GPU parity of the VV relation and critical-point classes against the CPU
oracle (restatement of src/cpu_extraction.cpp:259-289 and
src/alg/critPoints.cu:49-338).  Bar: bit-exact classes[3V]; VV rows equal as
sets (the reference leaves the order within a row unspecified; the device
emits ascending rows, so rows are compared after sorting the oracle's)."""
import numpy as np
import pytest

from oracle import binding as O

pytestmark = pytest.mark.gpu


# (variant, routing): vertex-lane form; bucketed form, bucketed owner form and
# bucketed form with the lean insert, each with one-level and two-level
# routing of the corner records
VARIANTS = [(1, 0), (2, 1), (2, 2), (3, 1), (3, 2), (4, 1), (4, 2)]
IDS = ["lane", "bucket-1level", "bucket-2level", "owner-1level", "owner-2level",
       "lean-1level", "lean-2level"]
BUCKETED, BUCKETED_IDS = VARIANTS[1:], IDS[1:]


def run_device(t, cells, n_points, scalars, cfg=None, twice=False, variant=0):
    with t.Mesh(cells, n_points, scalars) as m:
        if cfg:
            m.set_config(*cfg, variant=1)
        elif variant:
            v, route = variant if isinstance(variant, tuple) else (variant, 0)
            m.set_config(variant=v)
            if route:
                m.set_routing(route)
        m.vv_critpoints()
        if twice:
            m.vv_critpoints()
        off, vals = m.relation("VV")
        return off, vals[:, 0].copy(), m.classes()


def check(t, cells, n_points, scalars, cfg=None, twice=False, variant=0):
    print(f"case V={n_points} T={len(cells)} cfg={cfg} variant={variant}", flush=True)
    off, cols, cls = run_device(t, cells, n_points, scalars, cfg, twice, variant)
    VV = O.vv(cells, n_points).sorted_rows()
    assert np.array_equal(off, VV.off), "VV offsets differ"
    assert np.array_equal(cols, VV.val[:, 0]), "VV rows differ"
    ref = O.critpoints(O.vv(cells, n_points), np.asarray(scalars, np.float64), n_points)
    bad = np.nonzero((cls != ref).any(axis=1))[0]
    assert bad.size == 0, f"{bad.size} vertices differ, first {bad[:5]}: {cls[bad[:5]]} vs {ref[bad[:5]]}"
    return cls


@pytest.mark.parametrize("variant", VARIANTS, ids=IDS)
def test_config1_kuhn24(gpu, variant):
    m = gpu.kuhn(24, seed=1234)
    cls = check(gpu, m.cells, m.n_points, m.scalars, variant=variant)
    # uniform noise: every class occurs
    assert set(np.unique(cls[:, 2])) == {1, 2, 3, 4}


@pytest.mark.parametrize("cfg", [(32, 64), (32, 128), (32, 256), (64, 64), (64, 128)])
def test_lane_shapes_identical(gpu, cfg):
    m = gpu.kuhn(12, seed=7)
    check(gpu, m.cells, m.n_points, m.scalars, cfg=cfg)


@pytest.mark.parametrize("variant", VARIANTS, ids=IDS)
@pytest.mark.parametrize("order", ["path", "shuffled"])
def test_unsorted_cells(gpu, order, variant):
    m = gpu.kuhn(10, seed=3, order=order)
    check(gpu, m.cells, m.n_points, m.scalars, variant=variant)


def test_irregular_hubs_over_64_neighbours(gpu):
    # hubs push some vertex degrees past 64: serial kernels and the big fix path
    m = gpu.kuhn(12, seed=11, split_fraction=0.15, n_hubs=6, hub_splits=60, scalar="sines")
    VV = O.vv(m.cells, m.n_points)
    assert np.diff(VV.off).max() > 64
    for variant in VARIANTS:
        check(gpu, m.cells, m.n_points, m.scalars, variant=variant)
    check(gpu, m.cells, m.n_points, m.scalars, cfg=(64, 128))


@pytest.mark.parametrize("variant", VARIANTS, ids=IDS)
def test_all_ties_constant_field(gpu, variant):
    m = gpu.kuhn(8, seed=5, scalar="constant")
    check(gpu, m.cells, m.n_points, m.scalars, variant=variant)


@pytest.mark.parametrize("variant", VARIANTS, ids=IDS)
def test_linear_field_one_min_one_max(gpu, variant):
    m = gpu.kuhn(9, permute_vertices=False, permute_cells=True, jitter=False, scalar="linear")
    cls = check(gpu, m.cells, m.n_points, m.scalars, variant=variant)
    assert (cls[:, 2] == 1).sum() == 1 and (cls[:, 2] == 2).sum() == 1
    assert (cls[:, 2] == 3).sum() == m.n_points - 2


@pytest.mark.parametrize("variant", VARIANTS, ids=IDS)
def test_gaussian_field_f64_scalars(gpu, variant):
    m = gpu.kuhn(14, seed=21, scalar="gaussians")
    check(gpu, m.cells, m.n_points, m.scalars.astype(np.float64) * (1 + 1e-9), variant=variant)


@pytest.mark.parametrize("variant", VARIANTS, ids=IDS)
def test_nan_scalars(gpu, variant):
    m = gpu.kuhn(6, seed=2)
    s = m.scalars.copy()
    s[::7] = np.nan
    check(gpu, m.cells, m.n_points, s, variant=variant)


@pytest.mark.parametrize("variant", VARIANTS, ids=IDS)
def test_isolated_and_tiny(gpu, variant):
    # one tet plus three vertices no cell references: (0, 0, 4)
    cells = np.array([[0, 2, 4, 5]], np.int32)
    cls = check(gpu, cells, 8, np.arange(8, dtype=np.float32), variant=variant)
    assert cls[1].tolist() == [0, 0, 4] and cls[3].tolist() == [0, 0, 4]


@pytest.mark.parametrize("variant", VARIANTS, ids=IDS)
def test_empty_mesh(gpu, variant):
    off, cols, cls = run_device(gpu, np.zeros((0, 4), np.int32), 5, np.zeros(5, np.float32),
                                variant=variant)
    assert off.tolist() == [0] * 6 and cols.size == 0
    assert cls.tolist() == [[0, 0, 4]] * 5


@pytest.mark.parametrize("variant", VARIANTS, ids=IDS)
def test_repeat_is_deterministic(gpu, variant):
    m = gpu.kuhn(16, seed=9)
    a = run_device(gpu, m.cells, m.n_points, m.scalars, variant=variant)
    b = run_device(gpu, m.cells, m.n_points, m.scalars, twice=True, variant=variant)
    for x, y in zip(a, b):
        assert np.array_equal(x, y)


def test_invalid_cells_rejected(gpu):
    from tetragpu_amd._native import NativeError

    with pytest.raises(NativeError) as e:
        gpu.Mesh(np.array([[0, 1, 2, 9]], np.int32), 4, None)
    assert e.value.status == -4
    with pytest.raises(NativeError) as e:
        gpu.Mesh(np.array([[0, 1, 1, 2]], np.int32), 4, None)
    assert e.value.status == -4


def test_approx_max_vv(gpu):
    m = gpu.kuhn(10, seed=4)
    with gpu.Mesh(m.cells, m.n_points, m.scalars) as d:
        assert d.approx_max_vv() == O.approx_max_vv(m.cells, m.n_points) == 96


def test_labels6_match_oracle(gpu):
    m = gpu.kuhn(12, seed=13, scalar="gaussians")
    with gpu.Mesh(m.cells, m.n_points, m.scalars) as d:
        d.critpoints()
        lab = d.labels6()
        cls = d.classes()
    assert np.array_equal(lab, O.label6(cls))


@pytest.mark.parametrize("variant", VARIANTS, ids=IDS)
def test_hub_vertices_bucket_overflow(gpu, variant):
    # many vertices above 32 neighbours: the bucketed form routes them to its
    # serial kernel; both forms must agree with the oracle
    m = gpu.kuhn(10, seed=17, split_fraction=0.3, n_hubs=40, hub_splits=40, scalar="gaussians")
    check(gpu, m.cells, m.n_points, m.scalars, variant=variant)


@pytest.mark.parametrize("variant", BUCKETED, ids=BUCKETED_IDS)
def test_bucket_boundaries_partial_last_bucket(gpu, variant):
    # V not a multiple of the bucket size, unreferenced trailing vertices
    m = gpu.kuhn(7, seed=23)
    check(gpu, m.cells, m.n_points + 77, np.concatenate([m.scalars, np.zeros(77, np.float32)]),
          variant=variant)


@pytest.mark.parametrize("variant", BUCKETED, ids=BUCKETED_IDS)
def test_dense_bucket_over_capacity(gpu, variant):
    # one 128-vertex bucket receives ~25k corner records (more than the owner
    # form's per-bucket index list holds): every vertex takes the serial path
    cells, sc = soup(31, 200, 10000)
    check(gpu, cells, 200, sc, variant=variant)


def soup(seed, V, T):
    """Random tets (4 distinct vertices each) over V vertices: non-manifold,
    full of empty triangles, so the row-membership link rule merges
    components the shared-tet rule keeps apart."""
    rng = np.random.default_rng(seed)
    cells = np.array([rng.choice(V, 4, replace=False) for _ in range(T)], np.int32)
    sc = rng.integers(0, 8, V).astype(np.float32)  # many ties
    return cells, sc


@pytest.mark.parametrize("variant", VARIANTS, ids=IDS)
@pytest.mark.parametrize("seed,V,T", [(1, 30, 40), (2, 200, 300), (3, 60, 400),
                                      (4, 40, 1500), (5, 400, 700)])
def test_random_soup(gpu, variant, seed, V, T):
    cells, sc = soup(seed, V, T)
    check(gpu, cells, V, sc, variant=variant)


def test_non_flag_link_merges_components(gpu):
    # v=0 sits in two tets that share only v; tet (1, 4, 7, 8) joins neighbour
    # 1 of the first with neighbour 4 of the second without containing v
    cells = np.array([[0, 1, 2, 3], [0, 4, 5, 6], [1, 4, 7, 8]], np.int32)
    sc = np.array([10, 1, 2, 3, 4, 5, 6, 20, 20], np.float32)
    for variant in VARIANTS:
        cls = check(gpu, cells, 9, sc, variant=variant)
        # one lower component after the merge: a maximum, not a saddle
        assert cls[0].tolist() == [1, 0, 2]


def test_routing_auto_decision_and_parity(gpu):
    # grid-ordered vertices and cells: every 1024-tet tile spans few buckets
    # -> one level; with seeded vertex and cell permutations every tile spans
    # all buckets -> two levels.  Classes and VV agree across forced routings
    # and the vertex-lane form.
    for permuted, want in ((False, 1), (True, 2)):
        m = gpu.kuhn(56, seed=29, permute_vertices=permuted, permute_cells=permuted,
                     scalar="gaussians")
        assert m.n_points > 1024 * 128
        outs = []
        for form in ((1, 0), (2, 0), (2, 1), (2, 2)):
            with gpu.Mesh(m.cells, m.n_points, m.scalars) as d:
                d.set_config(variant=form[0])
                d.set_routing(form[1])
                d.vv_critpoints()
                if form == (2, 0):
                    assert d.routing() == want, permuted
                off, vals = d.relation("VV")
                outs.append((off, vals[:, 0].copy(), d.classes()))
        print(f"routing permuted={permuted} V={m.n_points} T={len(m.cells)}", flush=True)
        for o in outs[1:]:
            for x, y in zip(outs[0], o):
                assert np.array_equal(x, y)
