"""The serving loop: every request's result in a busy mix (several slots,
fusion of small requests, a tight HBM budget that makes requests wait) is
bit-identical to the same request served alone, and to the oracle; follow-up
VT/TT queries on resident relations equal the oracle's."""
import threading
import time

import numpy as np
import pytest

from oracle import binding as O

pytestmark = pytest.mark.gpu


def requests(seed, count):
    import tetragpu_amd as t
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        if i % 5 == 4:
            V = int(rng.integers(10, 50))
            cells = np.array([rng.choice(V, 4, replace=False) for _ in range(int(rng.integers(5, 60)))],
                             np.int32)
            out.append((cells, V, rng.integers(0, 4, V).astype(np.float32)))
        else:
            m = t.kuhn(int(rng.integers(2, 9)), seed=int(rng.integers(1 << 30)),
                       split_fraction=0.1 if i % 3 == 0 else 0.0, scalar="gaussians")
            out.append((m.cells, m.n_points, m.scalars))
    return out


def serve_all(reqs, followups, alone, **cfg):
    from tetragpu_amd.serve import Server
    results = []
    if alone:
        for cells, V, s in reqs:
            with Server(slots=1, fuse_max_meshes=1) as srv:
                rid = srv.submit(cells, V, s, followups)
                results.append(srv.wait(rid))
                srv.release(rid)
        return results, None
    with Server(**cfg) as srv:
        ids = []
        for k, (cells, V, s) in enumerate(reqs):
            ids.append(srv.submit(cells, V, s, followups))
            if k % 7 == 3:
                time.sleep(0.002)  # staggered arrivals
        results = []
        for i in ids:  # release as results arrive: the budget admits the rest
            results.append(srv.wait(i))
            srv.release(i)
        stats = srv.stats()
        assert stats["resident_bytes"] == 0
    return results, stats


def test_alone_equals_in_the_mix(gpu):
    from tetragpu_amd.serve import COPY, TT, VT
    reqs = requests(1, 30)
    print(f"serve {len(reqs)} requests", flush=True)
    alone, _ = serve_all(reqs, VT | TT | COPY, True)
    mixed, stats = serve_all(reqs, VT | TT | COPY, False, slots=3, hbm_budget_bytes=8 << 20,
                             fuse_max_cells=4000, fuse_max_meshes=8)
    assert stats["groups"] < len(reqs)  # some requests were fused
    assert any(r["group_size"] > 1 for r in mixed)
    for k, (a, b) in enumerate(zip(alone, mixed)):
        assert a["status"] == 0 and b["status"] == 0
        assert np.array_equal(a["classes"], b["classes"]), f"request {k} classes"
        for rel in ("VT", "TT"):
            assert np.array_equal(a[rel][0], b[rel][0]) and np.array_equal(a[rel][1], b[rel][1]), \
                f"request {k} {rel}"
    for k in (0, 4, len(reqs) - 1):
        cells, V, s = reqs[k]
        ref = O.critpoints(O.vv(cells, V), s.astype(np.float64), V)
        assert np.array_equal(mixed[k]["classes"], ref)
        VT_ref, TT_ref = O.vt(cells, V), O.tt(cells, V)
        assert np.array_equal(mixed[k]["VT"][0], VT_ref.off)
        assert np.array_equal(mixed[k]["VT"][1], VT_ref.val[:, 0])
        assert np.array_equal(mixed[k]["TT"][0], TT_ref.off)
        assert np.array_equal(mixed[k]["TT"][1], TT_ref.val[:, 0])


def test_followup_on_resident_relations(gpu):
    from tetragpu_amd.serve import TT, VT, Server
    reqs = requests(2, 6)
    with Server(slots=2, fuse_max_cells=10000, fuse_max_meshes=4) as srv:
        ids = [srv.submit(c, V, s) for c, V, s in reqs]
        first = [srv.wait(i) for i in ids]
        assert all("VT" not in r for r in first)
        for i, (cells, V, s) in zip(ids, reqs):
            TT_ref, VT_ref = O.tt(cells, V), O.vt(cells, V)
            # built and kept on the device: counts only
            r0 = srv.followup(i, VT | TT, copy=False)
            assert r0["tt_nnz"] == len(TT_ref.val) and r0["vt_nnz"] == len(VT_ref.val)
            assert r0["copy_ms"] == 0 and r0["followup_ms"] >= 0
            r = srv.followup(i, VT | TT)
            assert np.array_equal(r["TT"][0], TT_ref.off)
            assert np.array_equal(r["TT"][1], TT_ref.val[:, 0])
            assert np.array_equal(r["VT"][1], VT_ref.val[:, 0])
            assert r["upload_ms"] >= 0 and r["compute_ms"] >= 0
            srv.release(i)


def test_concurrent_submitters(gpu):
    from tetragpu_amd.serve import Server
    reqs = requests(3, 24)
    out = {}
    with Server(slots=2, hbm_budget_bytes=32 << 20, fuse_max_cells=3000, fuse_max_meshes=6) as srv:
        def client(part):
            for k in part:
                rid = srv.submit(*reqs[k])
                out[k] = srv.wait(rid)
                srv.release(rid)
        th = [threading.Thread(target=client, args=(range(j, len(reqs), 4),)) for j in range(4)]
        for x in th:
            x.start()
        for x in th:
            x.join()
    for k, (cells, V, s) in enumerate(reqs):
        ref = O.critpoints(O.vv(cells, V), s.astype(np.float64), V)
        assert np.array_equal(out[k]["classes"], ref), f"request {k}"


def test_degenerate_requests(gpu):
    """A mesh with no cells, one cell, and one cell beside an isolated vertex,
    served alone and fused into one group: classes equal the oracle's (a
    vertex in no cell has no lower or upper link: class 4)."""
    from tetragpu_amd.serve import Server
    one = np.array([[0, 1, 2, 3]], np.int32)
    reqs = [(np.zeros((0, 4), np.int32), 3, np.array([0.5, 0.25, 1.0], np.float32)),
            (one, 4, np.array([0.0, 1.0, 2.0, 3.0], np.float32)),
            (one, 5, np.array([3.0, 1.0, 2.0, 0.0, 9.0], np.float32))]
    for cfg in (dict(slots=1, fuse_max_meshes=1), dict(slots=1, fuse_max_meshes=8)):
        with Server(**cfg) as srv:
            ids = [srv.submit(c, V, s) for c, V, s in reqs]
            got = [srv.wait(i) for i in ids]
            for i in ids:
                srv.release(i)
        for k, ((cells, V, s), r) in enumerate(zip(reqs, got)):
            assert r["status"] == 0, (cfg, k)
            ref = O.critpoints(O.vv(cells, V), s.astype(np.float64), V)
            assert np.array_equal(r["classes"], ref), (cfg, k, r["classes"], ref)
    assert ref[4].tolist() == [0, 0, 4]
