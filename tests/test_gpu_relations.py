"""GPU parity of the id-assigning relations and those derived from them
against the CPU oracle.

Edge ids follow make_TE_and_VE (src/cpu_extraction.cpp:28-62), face ids
make_TF_and_VF (src/cpu_extraction.cpp:122-174), EV/ET/FV the transposes at
:95-120 and :225-243, VT :291-342.  FE/FT/EF/TT/FF/EE are named by flag only
in the reference (include/datatypes.h.in:173-178); their definitions are the
oracle's.  Bar: bit-exact ids, rows and row order."""
import numpy as np
import pytest

from oracle import binding as O

pytestmark = pytest.mark.gpu

ALL = ["TE", "EV", "ET", "TF", "FV", "FE", "FT", "EF", "VT", "TT", "FF", "EE", "VV", "VE", "VF"]


def oracle_all(cells, V):
    E, TE, VE = O.edges(cells, V)
    F, TF, VF = O.faces(cells, V)
    EV = O.ev(VE, E)
    FV = O.fv(VF, F)
    out = {"E": E, "F": F, "TE": TE, "VE": VE, "EV": EV, "ET": O.et(TE, E), "TF": TF,
           "VF": VF, "FV": FV, "FE": O.fe(FV, VE).reshape(F, 6), "VT": O.vt(cells, V),
           "TT": O.tt(cells, V), "EE": O.ee(EV, V)}
    out["EF"] = O.ef(out["FE"].reshape(F, 3, 2), E)
    try:
        FT = O.ft(TF, F)
        out["FT"] = FT
        out["FF"] = O.ff(TF, FT)
    except ValueError:
        out["FT"] = None
    return out


def dev_all(t, cells, V, names, variant=0):
    with t.Mesh(cells, V, None) as m:
        if variant:
            m.set_config(variant=variant)
        m.build(t.flags(*names))
        res = {"counts": m.counts()}
        for n in names:
            res[n] = m.relation(n)
        return res


def same_fixed(dev, ref, name):
    off, vals = dev
    assert off is None, name
    ref = np.asarray(ref).reshape(vals.shape[0], -1)
    assert vals.shape == ref.shape, f"{name} shape {vals.shape} vs {ref.shape}"
    bad = np.nonzero((vals != ref).any(axis=1))[0]
    assert bad.size == 0, f"{name}: {bad.size} rows differ, first {bad[:3]}: {vals[bad[:3]]} vs {ref[bad[:3]]}"


def same_csr(dev, ref, name):
    off, vals = dev
    assert np.array_equal(off, ref.off), f"{name} offsets differ"
    assert vals.shape == ref.val.shape, f"{name} shape {vals.shape} vs {ref.val.shape}"
    bad = np.nonzero((vals != ref.val).any(axis=1))[0]
    assert bad.size == 0, f"{name}: {bad.size} entries differ, first at {bad[:3]}"


def check_all(t, cells, V, variant=0, expect_ft=True):
    print(f"relations V={V} T={len(cells)} variant={variant}", flush=True)
    ref = oracle_all(cells, V)
    names = [n for n in ALL if n not in ("FT", "FF")] if ref["FT"] is None else ALL
    d = dev_all(t, cells, V, names, variant)
    assert d["counts"]["E"] == ref["E"] and d["counts"]["F"] == ref["F"]
    for n in ("TE", "EV", "TF", "FV", "FE"):
        same_fixed(d[n], ref[n], n)
    for n in ("VE", "VF", "ET", "EF", "VT", "TT", "EE"):
        same_csr(d[n], ref[n], n)
    if ref["FT"] is not None:
        same_fixed(d["FT"], ref["FT"], "FT")
        same_csr(d["FF"], ref["FF"], "FF")
    else:
        assert not expect_ft
    return ref


@pytest.mark.parametrize("variant", [1, 2])
def test_kuhn_sorted_cells_closed_forms(gpu, variant):
    n = 5
    m = gpu.kuhn(n, permute_vertices=False, permute_cells=False, jitter=False, order="sorted")
    ref = check_all(gpu, m.cells, m.n_points, variant)
    assert ref["E"] == 7 * n**3 + 9 * n**2 + 3 * n
    assert ref["F"] == 12 * n**3 + 6 * n**2


@pytest.mark.parametrize("order", ["path", "shuffled"])
def test_kuhn_permuted_unsorted(gpu, order):
    m = gpu.kuhn(6, seed=3, order=order)
    check_all(gpu, m.cells, m.n_points)


@pytest.mark.parametrize("hub_splits", [40, 90])
def test_kuhn_hubs(gpu, hub_splits):
    """High-valence hub vertices: rows beyond the LDS merge columns (100 words
    for TT and EE, 125 for ET, 48 for the edge walk) merge from global memory,
    and TT rows beyond 128 words take several chunks of the row pack."""
    m = gpu.kuhn(6, seed=11, split_fraction=0.2, n_hubs=4, hub_splits=hub_splits)
    ref = check_all(gpu, m.cells, m.n_points)
    vt_len = np.diff(ref["VT"].off)
    ev = np.asarray(ref["EV"]).reshape(-1, 2)
    deg = np.bincount(ev.ravel(), minlength=m.n_points)
    assert vt_len[m.cells].sum(axis=1).max() > 100
    assert np.diff(ref["TT"].off).max() > 128
    assert (vt_len[ev[:, 0]] + vt_len[ev[:, 1]]).max() > 125
    assert deg.max() > 48
    if hub_splits == 90:
        assert (deg[ev[:, 0]] + deg[ev[:, 1]]).max() > 100


@pytest.mark.parametrize("seed,V,T", [(1, 30, 40), (2, 200, 300), (3, 60, 400), (5, 400, 700)])
def test_random_soup(gpu, seed, V, T):
    rng = np.random.default_rng(seed)
    cells = np.array([rng.choice(V, 4, replace=False) for _ in range(T)], np.int32)
    check_all(gpu, cells, V, expect_ft=False)


def test_face_over_two_cells_rejected(gpu):
    # three cells on the triangle (0,1,2) at the same positions
    cells = np.array([[0, 1, 2, 3], [0, 1, 2, 4], [0, 1, 2, 5]], np.int32)
    from tetragpu_amd._native import NativeError

    with gpu.Mesh(cells, 6, None) as m:
        m.build(gpu.flags("TF", "FV"))
        with pytest.raises(NativeError) as e:
            m.build(gpu.flags("FT"))
        assert e.value.status == -4


def test_single_cube_fixture(gpu):
    # one Kuhn cube, cells in the generator's order: counts 19 edges, 18 faces
    m = gpu.kuhn(1, permute_vertices=False, permute_cells=False, jitter=False, order="sorted")
    ref = check_all(gpu, m.cells, m.n_points)
    assert ref["E"] == 19 and ref["F"] == 18


def test_empty_and_isolated(gpu):
    cells = np.zeros((0, 4), np.int32)
    d = dev_all(gpu, cells, 4, ["TE", "EV", "VE", "TF", "FV", "VF", "ET", "TT", "EE"])
    assert d["counts"]["E"] == 0 and d["counts"]["F"] == 0
    assert d["VE"][0].tolist() == [0] * 5 and d["VE"][1].shape[0] == 0
    cells = np.array([[3, 1, 6, 0]], np.int32)
    check_all(gpu, cells, 9)


@pytest.mark.skip(reason="this sequence (TE, then the vertex-lane hot path, then ET/faces) "
                  "made a GPU memory fault once, in a build before the id kernels were first "
                  "committed; no cause has been shown (the current and the first committed "
                  "kernels run it clean under AddressSanitizer on the CPU emulator, "
                  "tests/test_emu_kernels.py), so it stays off the GPU")
def test_relations_after_hot_path(gpu):
    # the vertex-lane VV form rewrites the corner lists; derived relations
    # built afterwards still see them sorted
    m = gpu.kuhn(5, seed=8)
    ref = oracle_all(m.cells, m.n_points)
    with gpu.Mesh(m.cells, m.n_points, m.scalars) as d:
        d.set_config(variant=1)
        d.build(gpu.flags("TE"))
        d.vv_critpoints()
        d.build(gpu.flags("ET", "TT", "VF"))
        same_csr(d.relation("ET"), ref["ET"], "ET")
        same_csr(d.relation("TT"), ref["TT"], "TT")
        same_csr(d.relation("VF"), ref["VF"], "VF")


@pytest.mark.parametrize("bv", [7, 8])
@pytest.mark.parametrize("route", [1, 2])
@pytest.mark.parametrize("case", ["dense", "kuhn", "partial"])
def test_vt_bucketed_routes(gpu, case, route, bv):
    # VT from the routed corner records: both routings, both bucket sizes,
    # buckets over the LDS capacity (rows staged in global memory) and a
    # partial last bucket with unreferenced vertices
    if case == "dense":
        rng = np.random.default_rng(31)
        V = 200
        cells = np.array([rng.choice(V, 4, replace=False) for _ in range(10000)], np.int32)
    elif case == "kuhn":
        m = gpu.kuhn(9, seed=4)
        cells, V = m.cells, m.n_points
    else:
        m = gpu.kuhn(5, seed=2)
        cells, V = m.cells, m.n_points + 77
    print(f"vt case={case} route={route} bv={bv}", flush=True)
    with gpu.Mesh(cells, V, None) as d:
        d.set_bucket(bv, 4096)
        d.set_routing(route)
        d.build(gpu.flags("VT"))
        same_csr(d.relation("VT"), O.vt(cells, V), "VT")
        assert d.approx_max_vv() == O.approx_max_vv(cells, V)


def test_duplicated_cell_faces_and_ff(gpu):
    # two cells on the same four vertices share all four faces: FT holds both,
    # and FF(f) is the other three faces once each
    cells = np.array([[0, 1, 2, 3], [0, 1, 2, 3], [4, 1, 2, 5]], np.int32)
    ref = check_all(gpu, cells, 6)
    assert ref["FT"] is not None
    assert [len(r) for r in np.split(ref["FF"].val[:, 0], ref["FF"].off[1:-1])][:4] == [3, 3, 3, 3]
