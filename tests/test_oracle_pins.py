"""Pins of the CPU oracle itself (no GPU): hand-derived fixtures, closed forms
of Kuhn grids, independent dictionary/BFS re-derivations in plain Python on
tiny meshes, and special cases that reduce to known answers.

Each check is chosen so that a plausible mistake in the restatement fails it:
sorting face keys (positional fixture, Kuhn unsorted counts), numbering edges
by key instead of first occurrence (TE fixture, dictionary re-derivation), a
reversed lower test (linear field puts the minimum at grid vertex 0), linking
lower neighbours by shared cells instead of row membership (non-flag merge),
dropping the id tie-break (constant field)."""
import itertools
import json
from pathlib import Path

import numpy as np
import pytest

import tetragpu_amd as t
from oracle import binding as O

GOLDEN = json.loads((Path(__file__).parent / "golden" / "hand_fixtures.json").read_text())
PAIRS = [(0, 1), (0, 2), (0, 3), (1, 2), (1, 3), (2, 3)]
SLOTS = [(0, 1, 2), (0, 1, 3), (0, 2, 3), (1, 2, 3)]


def rows(csr):
    return [csr.val[csr.off[i]:csr.off[i + 1]].tolist() for i in range(len(csr.off) - 1)]


def flat_rows(csr):
    return [[x[0] for x in r] for r in rows(csr)]


def all_relations(cells, V):
    E, TE, VE = O.edges(cells, V)
    F, TF, VF = O.faces(cells, V)
    EV, FV = O.ev(VE, E), O.fv(VF, F)
    FE = O.fe(FV, VE)
    out = dict(E=E, F=F, TE=TE, VE=VE, TF=TF, VF=VF, EV=EV, FV=FV, FE=FE,
               ET=O.et(TE, E), VV=O.vv(cells, V).sorted_rows(), VT=O.vt(cells, V),
               TT=O.tt(cells, V), EE=O.ee(EV, V), EF=O.ef(FE, E))
    try:
        out["FT"] = O.ft(TF, F)
        out["FF"] = O.ff(TF, out["FT"])
    except ValueError:
        out["FT"] = out["FF"] = None
    return out


# ---- hand-derived fixtures ------------------------------------------------------
@pytest.mark.parametrize("name", ["two_tets_sorted", "two_tets_unsorted"])
def test_hand_fixture(name):
    g = GOLDEN[name]
    cells, V = np.array(g["cells"]), g["n_points"]
    r = all_relations(cells, V)
    assert r["E"] == g["E"] and r["F"] == g["F"]
    assert r["TE"].tolist() == g["TE"] and r["TF"].tolist() == g["TF"]
    assert rows(r["VE"]) == g["VE"] and rows(r["VF"]) == g["VF"]
    assert r["EV"].tolist() == g["EV"] and r["FV"].tolist() == g["FV"]
    assert r["FT"].tolist() == g["FT"]
    for k in ("ET", "VV", "VT", "TT", "FF", "EE", "EF"):
        if k in g:
            assert flat_rows(r[k]) == g[k], k
    if "FE" in g:
        assert r["FE"].tolist() == g["FE"]


@pytest.mark.parametrize("case", range(2))
def test_hand_critpoints(case):
    g = GOLDEN["two_tets_sorted"]
    c = g["critpoints"][case]
    VV = O.vv(np.array(g["cells"]), g["n_points"])
    assert O.critpoints(VV, c["scalars"], g["n_points"]).tolist() == c["classes"]


# ---- Kuhn closed forms ------------------------------------------------------------
@pytest.mark.parametrize("n", [1, 2, 3, 4])
def test_kuhn_closed_forms(n):
    m = t.kuhn(n, permute_vertices=False, permute_cells=False, jitter=False)
    r = all_relations(m.cells, m.n_points)
    V, T, E, F = m.n_points, m.n_cells, r["E"], r["F"]
    assert (V, T) == ((n + 1) ** 3, 6 * n ** 3)
    assert E == 7 * n ** 3 + 9 * n ** 2 + 3 * n
    assert F == 12 * n ** 3 + 6 * n ** 2
    assert V - E + F - T == 1                      # a ball
    deg = np.diff(r["VV"].off)
    assert deg.sum() == 2 * E and deg.max() == (14 if n >= 2 else 7)
    if n >= 2:
        interior = [i + (n + 1) * (j + (n + 1) * k)
                    for i, j, k in itertools.product(range(1, n), repeat=3)]
        assert (deg[interior] == 14).all()
    assert r["VT"].val.shape[0] == 4 * T and r["ET"].val.shape[0] == 6 * T
    assert r["EF"].val.shape[0] == 3 * F
    boundary = (r["FT"][:, 1] < 0).sum()
    assert boundary == 12 * n ** 2 and (r["FT"][:, 0] >= 0).all()
    assert r["FF"].val.shape[0] == 6 * (F - boundary) + 3 * boundary
    assert O.approx_max_vv(m.cells, V) == (96 if n >= 2 else 32)
    # EV endpoints ascending and present in TE's cells
    assert (r["EV"][:, 0] < r["EV"][:, 1]).all()


def test_kuhn_unsorted_cells_split_faces():
    # positional keys: reordering corners inside cells can only split faces
    n = 3
    a = t.kuhn(n, permute_vertices=False, permute_cells=False, jitter=False)
    b = t.kuhn(n, seed=5, permute_vertices=False, permute_cells=False, jitter=False,
               order="shuffled")
    Fa = O.faces(a.cells, a.n_points)[0]
    Fb = O.faces(b.cells, b.n_points)[0]
    assert Fa == 12 * n ** 3 + 6 * n ** 2 and Fb > Fa
    # edges are canonical: the count does not change
    assert O.edges(a.cells, a.n_points)[0] == O.edges(b.cells, b.n_points)[0]


# ---- independent re-derivation on tiny meshes (plain Python) ----------------------
def soup(seed, V, T, sort=False):
    rng = np.random.default_rng(seed)
    cells = np.array([rng.choice(V, 4, replace=False) for _ in range(T)])
    return np.sort(cells, axis=1) if sort else cells


def py_edges_faces(cells, V):
    eid, te, ve = {}, [], [[] for _ in range(V)]
    for c in cells:
        row = []
        for i, j in PAIRS:
            key = (min(c[i], c[j]), max(c[i], c[j]))
            if key not in eid:
                eid[key] = len(eid)
                ve[key[0]].append([int(key[1]), eid[key]])
            row.append(eid[key])
        te.append(row)
    fid, tf, vf = {}, [], [[] for _ in range(V)]
    for c in cells:
        row = []
        for s in SLOTS:
            key = tuple(int(c[p]) for p in s)
            if key not in fid:
                fid[key] = len(fid)
                vf[key[0]].append([key[1], key[2], fid[key]])
            row.append(fid[key])
        tf.append(row)
    return eid, te, ve, fid, tf, vf


@pytest.mark.parametrize("seed,V,T,sort", [(1, 12, 20, False), (2, 30, 60, True),
                                           (3, 8, 40, False), (4, 50, 30, False)])
def test_first_occurrence_rederivation(seed, V, T, sort):
    cells = soup(seed, V, T, sort)
    eid, te, ve, fid, tf, vf = py_edges_faces(cells, V)
    r = all_relations(cells, V)
    assert r["E"] == len(eid) and r["F"] == len(fid)
    assert r["TE"].tolist() == te and r["TF"].tolist() == tf
    assert rows(r["VE"]) == ve and rows(r["VF"]) == vf
    inv = {i: k for k, i in eid.items()}
    assert [list(inv[i]) for i in range(len(eid))] == r["EV"].tolist()
    finv = {i: k for k, i in fid.items()}
    assert [[k[1], k[2], k[0]] for k in (finv[i] for i in range(len(fid)))] == r["FV"].tolist()
    # ET, VT, VV, TT, EE, FE, EF, FT, FF from their definitions
    vsets = [set(map(int, c)) for c in cells]
    assert flat_rows(r["ET"]) == [[c for c in range(T) if set(k) <= vsets[c]]
                                  for k in (inv[i] for i in range(len(eid)))]
    assert flat_rows(r["VT"]) == [[c for c in range(T) if v in vsets[c]] for v in range(V)]
    assert flat_rows(r["VV"]) == [sorted(set().union(*[vsets[c] for c in range(T)
                                                       if v in vsets[c]]) - {v})
                                  if any(v in s for s in vsets) else [] for v in range(V)]
    assert flat_rows(r["TT"]) == [[d for d in range(T) if d != c and vsets[c] & vsets[d]]
                                  for c in range(T)]
    assert flat_rows(r["EE"]) == [[j for j in range(len(eid)) if j != i and set(inv[i]) & set(inv[j])]
                                  for i in range(len(eid))]
    fe = []
    for i in range(len(fid)):
        lo, mid, hi = finv[i]
        fe.append([[max(a, b), eid[(min(a, b), max(a, b))]] for a, b in ((lo, mid), (lo, hi), (mid, hi))])
    assert r["FE"].tolist() == fe
    assert flat_rows(r["EF"]) == [[f for f in range(len(fid)) if any(x[1] == e for x in fe[f])]
                                  for e in range(len(eid))]
    holders = [[c for c in range(T) if i in tf[c]] for i in range(len(fid))]
    if max(len(h) for h in holders) <= 2:
        assert r["FT"].tolist() == [h + [-1] * (2 - len(h)) for h in holders]
        assert flat_rows(r["FF"]) == [sorted(set(f2 for c in holders[f] for f2 in tf[c]) - {f})
                                      for f in range(len(fid))]
    else:
        assert r["FT"] is None


# ---- critical points ------------------------------------------------------------------
def py_classes(cells, V, s):
    """Components by BFS over the row-membership graph, straight from the rule."""
    nb = [set() for _ in range(V)]
    for c in cells:
        for a in c:
            nb[a] |= set(map(int, c)) - {int(a)}
    out = []
    for v in range(V):
        low = {a for a in nb[v] if (s[a] < s[v]) or (s[a] == s[v] and a < v)}
        up = nb[v] - low
        counts = []
        for part in (low, up):
            seen, k = set(), 0
            for a in sorted(part):
                if a in seen:
                    continue
                k += 1
                stack = [a]
                seen.add(a)
                while stack:
                    x = stack.pop()
                    for y in nb[x] & part:
                        if y not in seen:
                            seen.add(y)
                            stack.append(y)
            counts.append(k)
        cl, cu = counts
        cls = 2 if (cl >= 1 and cu == 0) else 1 if (cl == 0 and cu >= 1) else 3 if (cl == 1 and cu == 1) else 4
        out.append([cl, cu, cls])
    return out


@pytest.mark.parametrize("seed,V,T", [(1, 12, 20), (2, 30, 50), (3, 9, 40), (7, 60, 40)])
def test_critpoints_bfs_rederivation(seed, V, T):
    cells = soup(seed, V, T)
    s = np.random.default_rng(seed + 100).integers(0, 4, V).astype(np.float64)  # ties
    got = O.critpoints(O.vv(cells, V), s, V).tolist()
    assert got == py_classes(cells, V, s)


def test_linear_field_one_min_one_max():
    n = 4
    m = t.kuhn(n, permute_vertices=False, permute_cells=True, jitter=False, scalar="linear")
    cls = O.critpoints(O.vv(m.cells, m.n_points), m.scalars, m.n_points)
    assert np.nonzero(cls[:, 2] == 1)[0].tolist() == [0]
    assert np.nonzero(cls[:, 2] == 2)[0].tolist() == [m.n_points - 1]
    assert (cls[:, 2] == 3).sum() == m.n_points - 2


def test_constant_field_reduces_to_ids():
    n = 3
    m = t.kuhn(n, permute_vertices=False, permute_cells=False, jitter=False, scalar="constant")
    VV = O.vv(m.cells, m.n_points)
    a = O.critpoints(VV, m.scalars, m.n_points)
    b = O.critpoints(VV, np.arange(m.n_points, dtype=np.float64), m.n_points)
    assert np.array_equal(a, b)
    assert (a[:, 2] == 1).sum() == 1 and (a[:, 2] == 2).sum() == 1


def test_non_flag_link_merges():
    cells = np.array([[0, 1, 2, 3], [0, 4, 5, 6], [1, 4, 7, 8]])
    s = np.array([10, 1, 2, 3, 4, 5, 6, 20, 20], np.float64)
    cls = O.critpoints(O.vv(cells, 9), s, 9)
    # the shared-cell link of vertex 0 has two lower components, the row rule one
    assert cls[0].tolist() == [1, 0, 2]


def test_isolated_and_nan():
    cells = np.array([[0, 1, 2, 3]])
    s = np.array([0.0, np.nan, 2.0, 3.0, 7.0])
    cls = O.critpoints(O.vv(cells, 5), s, 5)
    assert cls[4].tolist() == [0, 0, 4]                   # unreferenced vertex
    # NaN compares false both ways: vertex 1 sees every neighbour as upper,
    # and every other vertex sees vertex 1 as upper
    assert cls[1].tolist() == [0, 1, 1]
    assert cls[0].tolist() == [0, 1, 1]
    assert cls[3].tolist() == [1, 1, 3]


def test_label6_rules():
    cl = np.array([[0, 1, 1], [1, 0, 2], [1, 1, 3], [2, 1, 4], [1, 2, 4], [2, 2, 4],
                   [0, 0, 4], [3, 1, 4]], np.uint32)
    assert O.label6(cl).tolist() == [0, 3, 5, 1, 2, 4, 4, 4]


def test_omp_matches_serial():
    m = t.kuhn(7, seed=3, order="path", scalar="gaussians")
    VV, cls, _ = O.vv_cp_omp(m.cells, m.n_points, m.scalars, 4)
    ref_vv = O.vv(m.cells, m.n_points).sorted_rows()
    assert np.array_equal(VV.off, ref_vv.off) and np.array_equal(VV.val, ref_vv.val)
    assert np.array_equal(cls, O.critpoints(ref_vv, m.scalars.astype(np.float64), m.n_points))


# ---- partition loop (src/alg/critPoints.cu:654-1003, 1438-1491) -------------------------
def _soup(seed, V, T):
    rng = np.random.default_rng(seed)
    return np.array([rng.choice(V, 4, replace=False) for _ in range(T)])


def test_partitioned_one_partition_is_whole_mesh_but_voids():
    cells = np.array([[0, 1, 2, 3], [1, 2, 3, 5]])
    s = np.array([3.0, 1.0, 4.0, 1.0, 5.0, 9.0])
    whole = O.critpoints(O.vv(cells, 6), s, 6)
    part = O.critpoints_partitioned(cells, 6, s, np.zeros(6, np.int64))
    # vertex 4 is in no cell: the whole-mesh run classifies it (0, 0, 4),
    # the partition loop never sees it
    assert whole[4].tolist() == [0, 0, 4] and part[4].tolist() == [0, 0, 0]
    keep = np.arange(6) != 4
    assert np.array_equal(whole[keep], part[keep])


@pytest.mark.parametrize("n_parts", [2, 5, 17])
def test_partitioned_flag_mesh_equals_whole(n_parts):
    # Kuhn grids are flag complexes: any two neighbours of v that are joined
    # by an edge span a triangle with v, held by a cell of v's partition
    m = t.kuhn(6, seed=n_parts, scalar="gaussians")
    parts = np.random.default_rng(n_parts).integers(0, n_parts, m.n_points)
    s = m.scalars.astype(np.float64)
    whole = O.critpoints(O.vv(m.cells, m.n_points), s, m.n_points)
    assert np.array_equal(O.critpoints_partitioned(m.cells, m.n_points, s, parts), whole)


def test_partitioned_non_flag_known_answer():
    # the cell (1, 4, 7, 8) that joins vertex 0's two lower groups has no
    # vertex of partition 0, so partition 0 sees two lower components
    cells = np.array([[0, 1, 2, 3], [0, 4, 5, 6], [1, 4, 7, 8]])
    s = np.array([10, 1, 2, 3, 4, 5, 6, 20, 20], np.float64)
    parts = np.array([0, 1, 1, 1, 1, 1, 1, 1, 1])
    cls = O.critpoints_partitioned(cells, 9, s, parts)
    assert cls[0].tolist() == [2, 0, 2]
    # partition 1 holds every cell: its vertices match the whole mesh
    whole = O.critpoints(O.vv(cells, 9), s, 9)
    assert np.array_equal(cls[1:], whole[1:])
    # with vertex 1 or 4 in partition 0 the joining cell is included again
    parts[4] = 0
    assert O.critpoints_partitioned(cells, 9, s, parts)[0].tolist() == [1, 0, 2]


def test_partitioned_cell_lists_of_all_cells_equal_whole():
    cells = _soup(4, 30, 80)
    s = np.random.default_rng(4).integers(0, 6, 30).astype(np.float64)
    parts = np.arange(30) % 3
    full = np.ones((3, len(cells)), np.int32)
    whole = O.critpoints(O.vv(cells, 30), s, 30)
    got = O.critpoints_partitioned(cells, 30, s, parts, full)
    used = np.zeros(30, bool)
    used[np.unique(cells)] = True
    assert np.array_equal(got[used], whole[used]) and not got[~used].any()
