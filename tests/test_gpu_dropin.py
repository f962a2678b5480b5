"""The reference's driver and GPU builders through the C++ API on the GPU:
tetra_main builds every flagged relation with make_X_GPU (over the C ABI),
rebuilds them with the host builders and checks them with the reference's
tolerances and exactly, classifies critical points and exports the totals
(src/alg/critPoints.cu:1359-1514).  The host builders themselves are checked
against the oracle in test_dropin_cpu."""
import subprocess
from pathlib import Path

import numpy as np
import pytest

from oracle import binding as O

pytestmark = pytest.mark.gpu

ROOT = Path(__file__).resolve().parent.parent
MAIN = ROOT / "tetragpu_amd" / "lib" / "tetra_main"
ALL_FLAGS = ["--build_" + r for r in ("TE", "EV", "ET", "TF", "FV", "FE", "FT", "EF", "VT", "TT",
                                      "FF", "EE", "VV")]


def tetra_main(*args, timeout=240):
    print("tetra_main", *args, flush=True)
    return subprocess.run([str(MAIN), *map(str, args)], capture_output=True, text=True,
                          timeout=timeout)


def totals(text):
    out = {}
    for line in text.splitlines():
        if line.startswith("Total number of"):
            k, v = line.split(":")
            out[k.split()[-1]] = int(v)
    return out


@pytest.mark.parametrize("spec", ["kuhn:6:3:gaussians", "kuhn:5:11:uniform", "kuhn:4:2:constant"])
def test_all_relations_validated(gpu, spec, tmp_path):
    out = tmp_path / "classes.txt"
    r = tetra_main("-i", spec, "-e", out, "-d", "1", *ALL_FLAGS)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    assert "[validate] all checks passed" in r.stderr
    assert "FAILED" not in r.stderr
    # totals agree with the oracle on the same generated mesh
    import tetragpu_amd as t
    _, n, seed, scalar = spec.split(":")
    m = t.kuhn(int(n), seed=int(seed), scalar=scalar)
    cls = O.critpoints(O.vv(m.cells, m.n_points), m.scalars.astype(np.float64), m.n_points)
    tot = totals(out.read_text())
    assert tot["minima"] == (cls[:, 2] == 2).sum()   # the reference's label for class 2
    assert tot["maxima"] == (cls[:, 2] == 1).sum()
    assert tot["saddles"] == (cls[:, 2] == 4).sum()
    assert tot["regular"] == (cls[:, 2] == 3).sum()
    # each host-returning builder reports its device->host transfer and
    # translation (src/cuda_extraction.cu:326-336) in the reference's timer form
    for rel in ("TE", "EV", "ET", "TF", "FV", "FE", "FT", "EF", "VT", "TT", "FF", "EE", "VV"):
        for what in ("transfer", "translation"):
            line = f"Timer[{rel}_GPU] GPU Device->Host {what}: "
            got = [l for l in r.stdout.splitlines() if line in l]
            assert len(got) == 1, (line, r.stdout[-3000:])
            assert float(got[0].split(": ")[1]) >= 0


def test_vtu_input_default_flags(gpu, tmp_path):
    from test_dropin_cpu import write_vtu
    import tetragpu_amd as t

    m = t.kuhn(5, seed=8, scalar="sines")
    path = tmp_path / "m.vtu"
    write_vtu(path, m.cells, np.zeros((m.n_points, 3)), [("f", m.scalars)], "appended", True)
    r = tetra_main("-i", path, "-d", "0")   # defaults: VV + validate, classes to stdout
    assert r.returncode == 0, r.stderr[-2000:]
    cls = O.critpoints(O.vv(m.cells, m.n_points), m.scalars.astype(np.float64), m.n_points)
    assert totals(r.stdout)["saddles"] == (cls[:, 2] == 4).sum()


def test_bad_input_exits_with_failure(gpu, tmp_path):
    r = tetra_main("-i", tmp_path / "missing.vtu")
    assert r.returncode == 1 and "cannot open" in r.stderr


# ---- partition loop on the device (src/alg/critPoints.cu:654-1003, 1438-1491) -----------
from test_dropin_cpu import partition_cases, probe, probe_part  # noqa: E402,F401


@pytest.mark.parametrize("case", partition_cases(), ids=lambda c: c[0])
def test_gpu_partitioned_critpoints_equal_oracle(gpu, probe, tmp_path, case):  # noqa: F811
    name, cells, V, s, parts, pc = case
    print("partitioned", name, flush=True)
    got = probe_part(probe, tmp_path, cells, V, s, parts, "gpu", pc)
    assert np.array_equal(got, O.critpoints_partitioned(cells, V, s, parts, pc))


def test_tetra_main_partitions_from_vtu(gpu, tmp_path):
    from test_dropin_cpu import write_vtu
    import tetragpu_amd as t

    m = t.kuhn(6, seed=12, split_fraction=0.2, scalar="gaussians")
    parts = np.random.default_rng(5).integers(0, 4, m.n_points)
    path = tmp_path / "m.vtu"
    write_vtu(path, m.cells, np.zeros((m.n_points, 3)),
              [("f", m.scalars), ("part", (parts + 1).astype(np.int32))], "binary", True)
    out = tmp_path / "classes.txt"
    r = tetra_main("-i", path, "-a", "f", "-p", "part", "-e", out, "-d", "2", "--export_vertices")
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    assert "[validate] all checks passed" in r.stderr
    cls = O.critpoints_partitioned(m.cells, m.n_points, m.scalars.astype(np.float64), parts)
    text = out.read_text()
    tot = totals(text)
    assert tot["minima"] == (cls[:, 2] == 2).sum() and tot["maxima"] == (cls[:, 2] == 1).sum()
    assert tot["saddles"] == (cls[:, 2] == 4).sum() and tot["regular"] == (cls[:, 2] == 3).sum()
    for p in range(4):
        line = f"Partition {p} number of saddles:"
        got = [int(l.split(":")[1]) for l in text.splitlines() if l.startswith(line)]
        assert got == [int(((cls[:, 2] == 4) & (parts == p)).sum())]
    # partition 1's non-regular vertices, one line each (critPoints.cu:402-409)
    names = {1: "max", 2: "min", 4: "saddle"}
    want = [f"Partition 1 has a {names[int(c)]} for global vertex ID {v} "
            f"(upper: {cls[v, 0]}, lower: {cls[v, 1]})"
            for v, c in enumerate(cls[:, 2]) if parts[v] == 1 and c not in (0, 3)]
    assert want and [l for l in text.splitlines() if " has a " in l] == want
    # the per-vertex lines read back through pyclassify's rules
    from test_dropin_cpu import pyclassify_read
    got, insane = pyclassify_read(text)
    assert insane == [] and got[0] == []
    for k in (1, 2, 3, 4):
        assert got[k] == [int(v) for v in np.flatnonzero(cls[:, 2] == k)]


def perm_files(tmp_path, E, F, seed):
    rng = np.random.default_rng(seed)
    ep, fp = rng.permutation(E).astype(np.int64), rng.permutation(F).astype(np.int64)
    (tmp_path / "ep.i64").write_bytes(ep.tobytes())
    (tmp_path / "fp.i64").write_bytes(fp.tobytes())
    return ep, fp


def test_builders_honour_caller_numbering(gpu, probe, tmp_path):
    """make_{EV,TE,TF,FV,FE}_GPU(_return) given VE/VF whose ids are renumbered
    return their results in the caller's numbering, as the reference's
    builders do by computing from the uploaded tables
    (src/cuda_extraction.cu:85-252, 254-343, 491-844)."""
    import json
    from test_dropin_cpu import run

    m = gpu.kuhn(4, seed=9, order="shuffled")
    cells = np.ascontiguousarray(m.cells, np.int64)
    (tmp_path / "cells.i64").write_bytes(cells.tobytes())
    E, TE, VE = O.edges(m.cells, m.n_points)
    F, TF, VF = O.faces(m.cells, m.n_points)
    EV, FV = O.ev(VE, E), O.fv(VF, F)
    FE = O.fe(FV, VE).reshape(F, 6)
    for seed, identity in ((1, True), (2, False)):
        ep, fp = perm_files(tmp_path, E, F, seed)
        if identity:
            ep, fp = np.arange(E), np.arange(F)
            (tmp_path / "ep.i64").write_bytes(ep.astype(np.int64).tobytes())
            (tmp_path / "fp.i64").write_bytes(fp.astype(np.int64).tobytes())
        r = run(probe, "gpuids", tmp_path / "cells.i64", m.n_points, tmp_path / "ep.i64",
                tmp_path / "fp.i64")
        got = json.loads(r.stdout)
        ev = np.empty_like(EV); ev[ep] = EV
        fv = np.empty_like(FV); fv[fp] = FV
        fe = np.empty_like(FE); fe[fp] = FE
        fe[:, 1::2] = ep[fe[:, 1::2]]
        for k in ("", "_dev"):
            assert np.array_equal(got["EV" + k], ev)
            assert np.array_equal(got["TE" + k], ep[TE])
            assert np.array_equal(got["TF" + k], fp[TF])
            assert np.array_equal(got["FV" + k], fv)
            assert np.array_equal(got["FE" + k], fe)


def test_builders_reject_foreign_table(gpu, probe, tmp_path):
    from test_dropin_cpu import run

    m = gpu.kuhn(3, seed=4)
    (tmp_path / "cells.i64").write_bytes(np.ascontiguousarray(m.cells, np.int64).tobytes())
    E, _, _ = O.edges(m.cells, m.n_points)
    F, _, _ = O.faces(m.cells, m.n_points)
    perm_files(tmp_path, E, F, 3)
    r = run(probe, "gpuids", tmp_path / "cells.i64", m.n_points, tmp_path / "ep.i64",
            tmp_path / "fp.i64", "bad", check=False)
    assert r.returncode != 0 and "does not describe this mesh" in r.stderr
