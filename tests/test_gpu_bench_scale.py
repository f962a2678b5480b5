"""Parity at the benchmark workloads' full sizes.

- Kuhn 128^3 with the 32-Gaussian field (12,582,912 cells): every relation
  flag, VE/VF and the classes, each compared entry by entry with the oracle
  (first-occurrence ids of src/cpu_extraction.cpp:28-222, the transposes at
  :95-120 and :225-343, the flag-only relations as the oracle defines them),
  plus the Kuhn closed forms E = 7n^3+9n^2+3n, F = 12n^3+6n^2,
  nnz(ET) = 6T, nnz(VT) = 4T, nnz(EF) = 3F, nnz(FF) = 12T, 12n^2 boundary
  faces.
- The irregular-degree mesh (Kuhn 96^3, 15 % stellar splits, 32 hubs of 64
  splits; 7.7 M cells): VV, classes, TT and EE against the oracle.
- A batch of 256 Kuhn meshes, side log-uniform in [6, 40]: every mesh's VV
  rows and classes against the oracle run on that mesh alone.

Bar: bit-exact ids, rows, row order and classes.  Large arrays are compared
in slices so that no int64 copy of a device relation is made.
"""
import numpy as np
import pytest

from oracle import binding as O

pytestmark = pytest.mark.gpu

N2 = 128


def log(msg):
    print(msg, flush=True)


def same(a, b, name, chunk=1 << 26):
    a = np.asarray(a).reshape(len(a), -1)
    b = np.asarray(b).reshape(len(b), -1)
    assert a.shape == b.shape, f"{name}: shape {a.shape} vs {b.shape}"
    for i in range(0, a.shape[0], chunk):
        if not np.array_equal(a[i:i + chunk], b[i:i + chunk]):
            bad = np.nonzero((a[i:i + chunk] != b[i:i + chunk]).any(axis=1))[0]
            raise AssertionError(f"{name}: {bad.size} rows differ in slice {i}, first "
                                 f"{bad[:3] + i}")


def same_csr(dev, ref, name):
    off, vals = dev
    same(off, ref.off, name + " offsets")
    same(vals, ref.val, name)


@pytest.fixture(scope="module")
def cfg2(gpu):
    log("generating Kuhn 128^3, 32 gaussians")
    m = gpu.kuhn(N2, seed=1234, scalar="gaussians")
    d = gpu.Mesh(m.cells, m.n_points, m.scalars)
    log("building every relation flag + critical points on the device")
    d.build(gpu.ALL_FLAGS)
    d.vv_critpoints()
    d.sync()
    ref = {}
    yield m, d, ref
    d.close()


def oracle_ids(m, ref):
    if "E" not in ref:
        log("oracle: edge ids (serial first occurrence)")
        ref["E"], ref["TE"], ref["VE"] = O.edges(m.cells, m.n_points)
        ref["EV"] = O.ev(ref["VE"], ref["E"])
    if "F" not in ref:
        log("oracle: face ids (serial first occurrence)")
        ref["F"], ref["TF"], ref["VF"] = O.faces(m.cells, m.n_points)
        ref["FV"] = O.fv(ref["VF"], ref["F"])
    return ref


@pytest.mark.timeout(900)
def test_cfg2_vv_vt_classes(cfg2):
    m, d, _ = cfg2
    T, V = m.n_cells, m.n_points
    log("oracle: VV + classes (OpenMP form), VT")
    VV, cls, _ = O.vv_cp_omp(m.cells, V, m.scalars)
    same_csr(d.relation("VV"), VV, "VV")
    same(d.classes(), cls, "classes")
    VT = O.vt(m.cells, V)
    assert VT.val.shape[0] == 4 * T
    same_csr(d.relation("VT"), VT, "VT")


@pytest.mark.timeout(900)
def test_cfg2_edges(cfg2):
    m, d, ref = cfg2
    n, T = N2, m.n_cells
    oracle_ids(m, ref)
    E = ref["E"]
    assert E == 7 * n**3 + 9 * n**2 + 3 * n
    assert d.counts()["E"] == E
    same(d.relation("TE")[1], ref["TE"], "TE")
    same(d.relation("EV")[1], ref["EV"], "EV")
    same_csr(d.relation("VE"), ref["VE"], "VE")
    log("oracle: ET")
    ET = O.et(ref["TE"], E)
    assert ET.val.shape[0] == 6 * T
    same_csr(d.relation("ET"), ET, "ET")


@pytest.mark.timeout(900)
def test_cfg2_faces(cfg2):
    m, d, ref = cfg2
    n, T = N2, m.n_cells
    oracle_ids(m, ref)
    E, F = ref["E"], ref["F"]
    assert F == 12 * n**3 + 6 * n**2
    assert d.counts()["F"] == F
    same(d.relation("TF")[1], ref["TF"], "TF")
    same(d.relation("FV")[1], ref["FV"], "FV")
    same_csr(d.relation("VF"), ref["VF"], "VF")
    log("oracle: FE, FT, EF, FF")
    FE = O.fe(ref["FV"], ref["VE"])
    same(d.relation("FE")[1], FE.reshape(F, 6), "FE")
    FT = O.ft(ref["TF"], F)
    same(d.relation("FT")[1], FT, "FT")
    EF = O.ef(FE, E)
    assert EF.val.shape[0] == 3 * F
    same_csr(d.relation("EF"), EF, "EF")
    FF = O.ff(ref["TF"], FT)
    # interior faces see 6 others, boundary faces 3: nnz = 6 F_in + 3 F_b = 12 T
    assert int((FT[:, 1] < 0).sum()) == 12 * n**2 and FF.val.shape[0] == 12 * T
    same_csr(d.relation("FF"), FF, "FF")


@pytest.mark.timeout(900)
def test_cfg2_tt(cfg2):
    m, d, _ = cfg2
    log("oracle: TT (row-parallel)")
    TT = O.tt(m.cells, m.n_points)
    log(f"TT nnz {TT.val.shape[0]}")
    same_csr(d.relation("TT"), TT, "TT")


@pytest.mark.timeout(900)
def test_cfg2_ee(cfg2):
    m, d, ref = cfg2
    oracle_ids(m, ref)
    log("oracle: EE (row-parallel)")
    EE = O.ee(ref["EV"], m.n_points)
    same_csr(d.relation("EE"), EE, "EE")


@pytest.mark.timeout(900)
def test_cfg3_irregular(gpu):
    log("generating the irregular-degree mesh (Kuhn 96^3, 15 % splits, 32 hubs x 64)")
    m = gpu.kuhn(96, seed=5, split_fraction=0.15, n_hubs=32, hub_splits=64, scalar="sines")
    V = m.n_points
    with gpu.Mesh(m.cells, V, m.scalars) as d:
        d.vv_critpoints()
        d.build(gpu.flags("TT", "EE"))
        d.sync()
        log("oracle: VV + classes")
        VV, cls, _ = O.vv_cp_omp(m.cells, V, m.scalars)
        assert np.diff(VV.off).max() > 60
        same_csr(d.relation("VV"), VV, "VV")
        same(d.classes(), cls, "classes")
        log("oracle: TT")
        same_csr(d.relation("TT"), O.tt(m.cells, V), "TT")
        log("oracle: edges, EE")
        E, _, VE = O.edges(m.cells, V)
        same_csr(d.relation("EE"), O.ee(O.ev(VE, E), V), "EE")


@pytest.mark.timeout(900)
def test_cfg4_batch_256(gpu):
    rng = np.random.default_rng(4096)
    sides = np.floor(np.exp(rng.uniform(np.log(6), np.log(41), 256))).astype(int)
    meshes = [gpu.kuhn(int(n), seed=7 * 100003 + i, scalar="uniform") for i, n in enumerate(sides)]
    log(f"batch of {len(meshes)} meshes, {sum(x.n_cells for x in meshes)} cells, "
        f"sides {sides.min()}..{sides.max()}")
    with gpu.Batch([(x.cells, x.n_points, x.scalars) for x in meshes]) as b:
        b.vv_critpoints()
        b.sync()
        cls = b.classes()
        for i, x in enumerate(meshes):
            VV, ref, _ = O.vv_cp_omp(x.cells, x.n_points, x.scalars)
            assert np.array_equal(b.mesh_classes(i, cls), ref), f"mesh {i} classes"
            off, vals = b.vv(i)
            assert np.array_equal(off, VV.off) and np.array_equal(vals, VV.val[:, 0]), \
                f"mesh {i} VV"


@pytest.mark.timeout(900)
def test_above_2_24_vertices_vertex_lane_form(gpu):
    """Kuhn 256^3 (100,663,296 cells, 16,974,593 > 2^24 vertices): the
    bucketed form packs 24-bit ids, so the automatic choice takes the
    vertex-lane form here.  VV and classes against the OpenMP oracle."""
    import os
    import time

    n = 256
    m = gpu.kuhn(n, seed=7, scalar="gaussians")
    assert m.n_points > 1 << 24
    log(f"kuhn{n} V={m.n_points} T={m.n_cells}")
    with gpu.Mesh(m.cells, m.n_points, m.scalars) as d:
        d.vv_critpoints()
        d.sync()
        t0 = time.perf_counter()
        d.vv_critpoints()
        d.sync()
        log(f"vertex-lane form: {(time.perf_counter() - t0) * 1e3:.1f} ms per run")
        off, vals = d.relation("VV")
        cls = d.classes()
    E = 7 * n**3 + 9 * n**2 + 3 * n
    assert off[-1] == 2 * E
    VV, ref, _ = O.vv_cp_omp(m.cells, m.n_points, m.scalars, min(16, os.cpu_count() or 1))
    same(off, VV.off, "VV offsets")
    same(vals[:, 0], VV.val[:, 0], "VV columns")
    same(cls, ref, "classes")
