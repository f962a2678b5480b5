"""The C++ API with the reference's names (include/tetra/*.h), host side:
CPU builders and critical points against the oracle, argparse flag
semantics (src/argparse.cu:52-326), export_classes output
(src/alg/critPoints.cu:340-453) and the .vtu reader (formats written here by
a small writer of the VTK XML layout)."""
import base64
import json
import os
import struct
import subprocess
import zlib
from pathlib import Path

import numpy as np
import pytest

import tetragpu_amd as t
from oracle import binding as O

ROOT = Path(__file__).resolve().parent.parent
LIB = ROOT / "tetragpu_amd" / "lib"


@pytest.fixture(scope="session")
def probe(tmp_path_factory):
    exe = tmp_path_factory.mktemp("probe") / "dropin_probe"
    cmd = ["g++", "-O1", "-std=c++17", f"-I{ROOT / 'include' / 'tetra'}", f"-I{ROOT / 'include'}",
           str(ROOT / "tests" / "cpp" / "dropin_probe.cpp"), "-o", str(exe), f"-L{LIB}",
           "-ltetra_host", "-ltetra_hip", "-L/opt/rocm/lib", "-lamdhip64",
           f"-Wl,-rpath,{LIB}", "-Wl,-rpath,/opt/rocm/lib"]
    subprocess.run(cmd, check=True, capture_output=True, text=True)
    return exe


def run(probe, *args, check=True):
    r = subprocess.run([str(probe), *map(str, args)], capture_output=True, text=True, timeout=120)
    if check:
        assert r.returncode == 0, r.stderr
    return r


def cpu_json(probe, tmp_path, cells, V, scalars=None):
    cp = tmp_path / "cells.i64"
    np.ascontiguousarray(cells, np.int64).tofile(cp)
    args = ["cpu", cp, V]
    if scalars is not None:
        sp = tmp_path / "s.f64"
        np.ascontiguousarray(scalars, np.float64).tofile(sp)
        args.append(sp)
    return json.loads(run(probe, *args).stdout)


def last_json(r):
    return json.loads(r.stdout.strip().splitlines()[-1])


def csr_rows(c):
    return [c.val[c.off[i]:c.off[i + 1]].reshape(-1).tolist() for i in range(len(c.off) - 1)]


def meshes():
    rng = np.random.default_rng(3)
    soup = np.array([rng.choice(40, 4, replace=False) for _ in range(90)])
    k = t.kuhn(3, seed=9, order="shuffled")
    k2 = t.kuhn(4, seed=2)
    return [("soup", soup, 40), ("kuhn3_shuffled", k.cells, k.n_points),
            ("kuhn4", k2.cells, k2.n_points)]


@pytest.mark.parametrize("name,cells,V", meshes(), ids=lambda x: x if isinstance(x, str) else "")
def test_cpu_builders_equal_oracle(probe, tmp_path, name, cells, V):
    j = cpu_json(probe, tmp_path, cells, V)
    E, TE, VE = O.edges(cells, V)
    F, TF, VF = O.faces(cells, V)
    EV, FV = O.ev(VE, E), O.fv(VF, F)
    FE = O.fe(FV, VE)
    assert j["E"] == E and j["F"] == F and j["same_counts"] == 1
    assert j["TE"] == TE.tolist() and j["TF"] == TF.tolist()
    assert j["VE"] == csr_rows(VE) and j["VF"] == csr_rows(VF)
    assert j["EV"] == EV.tolist() and j["FV"] == FV.tolist()
    assert j["FE"] == FE.reshape(F, 6).tolist()
    assert j["ET"] == csr_rows(O.et(TE, E)) and j["EF"] == csr_rows(O.ef(FE, E))
    assert j["VV"] == csr_rows(O.vv(cells, V))          # insertion order, exactly
    assert j["VT"] == csr_rows(O.vt(cells, V)) and j["TT"] == csr_rows(O.tt(cells, V))
    assert j["EE"] == csr_rows(O.ee(EV, V))
    try:
        FT = O.ft(TF, F)
    except ValueError:
        FT = None
    if FT is not None:
        assert j["FT"] == FT.tolist() and j["FF"] == csr_rows(O.ff(TF, FT))
    else:
        assert "FT" not in j


@pytest.mark.parametrize("seed", [1, 2])
def test_cpu_critpoints_equal_oracle(probe, tmp_path, seed):
    m = t.kuhn(5, seed=seed, scalar="gaussians")
    s = m.scalars.astype(np.float64)
    s[1::11] = s[0]  # ties with vertex 0
    j = cpu_json(probe, tmp_path, m.cells, m.n_points, s)
    ref = O.critpoints(O.vv(m.cells, m.n_points), s, m.n_points)
    assert np.array_equal(np.array(j["classes"]).reshape(-1, 3), ref)


# ---- argparse ----------------------------------------------------------------------
def parse(probe, *args):
    return run(probe, "parse", *args)


def test_parse_defaults(probe):
    j = json.loads(parse(probe, "-g", "1").stdout)
    assert j["flags"] == "10000000000001" and j["fileName"] == "tiny.vtu"
    assert j["threads"] == 1 and j["max_VV"] == -1 and j["no_partitioning"] == 1


def test_parse_flags_only_switch_on(probe):
    j = json.loads(parse(probe, "-g", "1", "--build_TE", "--build_VT", "--build_EE").stdout)
    assert j["flags"] == "11001000000011"
    j = json.loads(parse(probe, "-g", "1", "--no_validate").stdout)
    assert j["flags"] == "10000000000000"


def test_parse_values(probe):
    j = json.loads(parse(probe, "-g", "1", "-i", "m.vtu", "-t", "8", "-e", "out.txt", "-a",
                         "f", "-v", "96", "-d", "0").stdout)
    assert (j["fileName"], j["threads"], j["export"], j["arrayname"], j["max_VV"], j["debug"]) == \
        ("m.vtu", 8, "out.txt", "f", 96, 0)


@pytest.mark.parametrize("bad", [["-t", "0"], ["-v", "0"], ["-d", "3"], ["--bogus"], ["-g", "-2"]])
def test_parse_rejects(probe, bad):
    r = run(probe, "parse", "-g", "1", *bad, check=False)
    assert r.returncode == 1


def test_parse_help(probe):
    r = run(probe, "parse", "-h")
    assert "--build_VV" in r.stdout and "--input" in r.stdout


# ---- export_classes -----------------------------------------------------------------
def test_export_classes_labels_and_insanity(probe, tmp_path):
    cls = np.array([[0, 1, 1], [1, 0, 2], [1, 0, 2], [1, 1, 3], [2, 1, 4], [0, 0, 4], [0, 0, 0],
                    [1, 0, 3]], np.uint32)
    p = tmp_path / "c.u32"
    cls.tofile(p)
    r = run(probe, "export", p, len(cls), 2)
    lines = r.stdout.splitlines()
    # the reference's labels: "minima" counts class 2, "maxima" class 1
    assert lines[0].split() == ["Total", "number", "of", "minima:", "2"]
    assert lines[1].split()[-1] == "1" and lines[2].split()[-1] == "2"
    assert lines[3].split()[-1] == "2" and lines[4].split()[-1] == "1"
    assert any(l.startswith("Partition 0 number of minima:") for l in lines)
    assert "Insanity detected" in r.stderr and " 1 points" in r.stderr  # (1,0,3)


def pyclassify_read(text):
    """The line rules of the reference's pyclassify.py:37-59, restated: lines
    opening with 'Total' or 'Partition' are skipped; a line holding 'Upper:'
    and 'Lower:' is the long form (id = third word, class = first class name
    found in it); any other line ends in its class number, and its vertex id
    is the second-last word (INSANITY lines are set aside)."""
    names = ["NULL", "min", "max", "regular", "saddle"]
    classes, insane = {k: [] for k in range(5)}, []
    for l in text.splitlines(keepends=True):
        if l.startswith("Total") or l.startswith("Partition"):
            continue
        if "Upper:" in l and "Lower:" in l:
            vid = int(l.split(" ", 3)[2])
            for k, n in enumerate(names):
                if n in l:
                    classes[k].append(vid)
                    break
            continue
        last = int(l.rstrip().rsplit(" ", 1)[1])
        if "INSANITY" in l:
            insane.append(last)
        else:
            classes[last].append(int(l.rsplit(" ", 3)[1]))
    return classes, insane


def test_export_partition_detail_and_vertex_lines(probe, tmp_path):
    # (upper, lower, class) per vertex; partitions 0/1 alternate in pairs
    cls = np.array([[0, 1, 1], [1, 0, 2], [2, 2, 4], [1, 1, 3], [0, 2, 1], [1, 1, 3],
                    [0, 0, 0], [3, 2, 4]], np.uint32)
    parts = np.array([0, 1, 1, 0, 1, 1, 0, 1], np.int32)
    cp, pp = tmp_path / "c.u32", tmp_path / "p.i32"
    cls.tofile(cp)
    parts.tofile(pp)
    r = run(probe, "export", cp, len(cls), 0, f"parts={pp}:2")
    detail = [l for l in r.stdout.splitlines() if " has a " in l]
    # critPoints.cu:402-409: partition 1, class neither void nor regular, in
    # vertex order, before the totals
    assert detail == [
        "Partition 1 has a min for global vertex ID 1 (upper: 1, lower: 0)",
        "Partition 1 has a saddle for global vertex ID 2 (upper: 2, lower: 2)",
        "Partition 1 has a max for global vertex ID 4 (upper: 0, lower: 2)",
        "Partition 1 has a saddle for global vertex ID 7 (upper: 3, lower: 2)"]
    assert r.stdout.splitlines()[len(detail)].startswith("Total number of minima:")
    assert not any(l.startswith("A ") for l in r.stdout.splitlines())
    # without partition ids nothing is partition 1
    r0 = run(probe, "export", cp, len(cls), 0)
    assert " has a " not in r0.stdout

    r = run(probe, "export", cp, len(cls), 2, f"parts={pp}:2", "vertices")
    got, insane = pyclassify_read(r.stdout)
    assert insane == []
    want = {k: [v for v in range(len(cls)) if cls[v, 2] == k] for k in range(5)}
    want[0] = []  # void vertices get no line
    assert got == want


# ---- .vtu reader ----------------------------------------------------------------------
VTK_TYPES = {np.dtype(np.int32): "Int32", np.dtype(np.int64): "Int64",
             np.dtype(np.float32): "Float32", np.dtype(np.float64): "Float64",
             np.dtype(np.uint8): "UInt8"}


def write_vtu(path, cells, pts, arrays, fmt="ascii", compress=False, header="UInt32",
              encoding="raw", cell_arrays=()):
    hdt = "I" if header == "UInt32" else "Q"
    appended = bytearray()

    def blob(a):
        raw = a.tobytes()
        if not compress:
            return struct.pack("<" + hdt, len(raw)), raw
        bs = 1 << 10
        blocks = [raw[i:i + bs] for i in range(0, len(raw), bs)] or [b""]
        comp = [zlib.compress(b) for b in blocks]
        last = len(blocks[-1]) if len(blocks[-1]) != bs else 0
        h = struct.pack("<" + hdt * (3 + len(comp)), len(comp), bs, last, *map(len, comp))
        return h, b"".join(comp)

    def data_array(name, a, ncomp=1):
        a = np.ascontiguousarray(a)
        attrs = f'type="{VTK_TYPES[a.dtype]}" Name="{name}" NumberOfComponents="{ncomp}"'
        if fmt == "ascii":
            return f'<DataArray {attrs} format="ascii">{" ".join(map(repr, a.reshape(-1).tolist()))}</DataArray>'
        h, d = blob(a)
        if fmt == "binary":
            if compress:
                txt = base64.b64encode(h).decode() + base64.b64encode(d).decode()
            else:
                txt = base64.b64encode(h + d).decode()
            return f'<DataArray {attrs} format="binary">{txt}</DataArray>'
        off = len(appended)
        if encoding == "raw":
            appended.extend(h + d)
        else:
            appended.extend(base64.b64encode(h) + base64.b64encode(d) if compress
                            else base64.b64encode(h + d))
        return f'<DataArray {attrs} format="appended" offset="{off}"/>'

    T, V = len(cells), len(pts)
    pd = "".join(data_array(n, a) for n, a in arrays)
    cd = "".join(data_array(n, a) for n, a in cell_arrays)
    cd = f"<CellData>{cd}</CellData>" if cd else ""
    body = (f'<Piece NumberOfPoints="{V}" NumberOfCells="{T}"><PointData>{pd}</PointData>{cd}'
            f'<Points>{data_array("Points", pts.astype(np.float32), 3)}</Points><Cells>'
            f'{data_array("connectivity", cells.astype(np.int64).reshape(-1))}'
            f'{data_array("offsets", (4 * np.arange(1, T + 1)).astype(np.int64))}'
            f'{data_array("types", np.full(T, 10, np.uint8))}</Cells></Piece>')
    comp = ' compressor="vtkZLibDataCompressor"' if compress else ""
    head = (f'<?xml version="1.0"?>\n<VTKFile type="UnstructuredGrid" version="1.0" '
            f'byte_order="LittleEndian" header_type="{header}"{comp}><UnstructuredGrid>{body}'
            f'</UnstructuredGrid>').encode()
    tail = b"</VTKFile>\n"
    if fmt == "appended":
        tail = (f'<AppendedData encoding="{encoding}">_'.encode() + bytes(appended)
                + b"</AppendedData></VTKFile>\n")
    Path(path).write_bytes(head + tail)


@pytest.mark.parametrize("fmt,compress,header,encoding", [
    ("ascii", False, "UInt32", "raw"), ("binary", False, "UInt32", "raw"),
    ("binary", True, "UInt64", "raw"), ("appended", False, "UInt64", "raw"),
    ("appended", True, "UInt32", "raw"), ("appended", False, "UInt32", "base64"),
    ("appended", True, "UInt64", "base64")])
def test_vtu_reader(probe, tmp_path, fmt, compress, header, encoding):
    m = t.kuhn(3, seed=4, scalar="gaussians")
    pts = np.random.default_rng(0).random((m.n_points, 3))
    second = np.arange(m.n_points, dtype=np.float64) * 0.5
    path = tmp_path / "m.vtu"
    write_vtu(path, m.cells, pts, [("first", m.scalars.astype(np.float32)), ("second", second)],
              fmt, compress, header, encoding)
    j = last_json(run(probe, "load", path))
    assert j["n_points"] == m.n_points and j["n_cells"] == m.n_cells
    assert j["cells"] == m.cells.reshape(-1).tolist()
    assert np.array_equal(np.array(j["scalars"]), m.scalars.astype(np.float64))
    j2 = last_json(run(probe, "load", path, "second"))
    assert np.array_equal(np.array(j2["scalars"]), second)


def test_vtu_rejects_non_tets(probe, tmp_path):
    path = tmp_path / "bad.vtu"
    path.write_text('<VTKFile type="UnstructuredGrid"><UnstructuredGrid><Piece NumberOfPoints="4" '
                    'NumberOfCells="1"><Cells><DataArray type="Int64" Name="connectivity" '
                    'format="ascii">0 1 2</DataArray><DataArray type="Int64" Name="offsets" '
                    'format="ascii">3</DataArray></Cells></Piece></UnstructuredGrid></VTKFile>')
    r = run(probe, "load", path, check=False)
    assert r.returncode == 1 and "tetrahedra" in r.stderr


def test_kuhn_spec_input(probe):
    j = last_json(run(probe, "load", "kuhn:3:7:linear"))
    m = t.kuhn(3, seed=7, permute_vertices=False, jitter=False, scalar="linear")
    assert j["cells"] == m.cells.reshape(-1).tolist()
    assert j["scalars"] == m.scalars.astype(np.float64).tolist()


# ---- partitions (src/vtk_load.cu:148-204, src/alg/critPoints.cu:654-1003) -------------
@pytest.mark.parametrize("raw,want", [
    ([1, 2, 2, 3, 1], [0, 1, 1, 2, 0]),        # 1-based ids are lowered
    ([1, 0, 2, 1], [1, 0, 2, 1]),              # a 0 switches to 0-based, earlier ids raised back
    ([2, 0, 1, 0], [2, 0, 1, 0]),
    ([1, 0, -1, 2], [1, 1, 0, 2]),             # a later negative value moves the switch point
    ([1.9, 2.2, 1.0], [0, 1, 0]),              # fractions truncate
])
def test_vtu_partition_reindexing(probe, tmp_path, raw, want):
    V = len(raw)
    cells = np.array([[0, 1, 2, 3]]) if V >= 4 else np.array([[0, 1, 2, 0]])[:0]
    path = tmp_path / "p.vtu"
    write_vtu(path, cells, np.zeros((V, 3)), [("s", np.arange(V, dtype=np.float64)),
                                               ("part", np.array(raw, np.float64))])
    j = last_json(run(probe, "load", path, "s", "part"))
    assert j["partition_ids"] == want
    assert j["n_partitions"] == len(set(want))
    assert j["n_per_partition"] == [want.count(p) for p in range(len(set(want)))]
    assert j["partition_cells"] == []


def test_vtu_partition_errors(probe, tmp_path):
    path = tmp_path / "p.vtu"
    write_vtu(path, np.array([[0, 1, 2, 3]]), np.zeros((4, 3)),
              [("s", np.arange(4.0)), ("part", np.array([1.0, 1.0, 6.0, 6.0]))])
    r = run(probe, "load", path, "s", "part", check=False)
    assert r.returncode == 1 and "not 0..1" in r.stderr
    r = run(probe, "load", path, "s", "missing", check=False)
    assert r.returncode == 1 and "Unable to find partitioning name 'missing'" in r.stderr


def test_vtu_default_single_partition(probe, tmp_path):
    j = last_json(run(probe, "load", "kuhn:2:1"))
    assert j["n_partitions"] == 1 and j["partition_ids"] == [0] * j["n_points"]
    assert j["n_per_partition"] == [j["n_points"]]


def test_vtu_partition_cell_lists(probe, tmp_path):
    m = t.kuhn(2, seed=3)
    parts = (np.arange(m.n_points) % 2).astype(np.int32)
    rng = np.random.default_rng(1)
    lists = [rng.integers(0, 2, m.n_cells).astype(np.int32) for _ in range(2)]
    path = tmp_path / "p.vtu"
    write_vtu(path, m.cells, np.zeros((m.n_points, 3)),
              [("s", m.scalars.astype(np.float64)), ("part", parts + 1)],
              cell_arrays=[("partition_cells_1", lists[1]), ("partition_cells_0", lists[0])])
    j = last_json(run(probe, "load", path, "s", "part"))
    assert j["partition_ids"] == parts.tolist()
    assert j["partition_cells"] == np.concatenate(lists).tolist()
    # one list missing: the vertex rule applies (no lists kept)
    write_vtu(path, m.cells, np.zeros((m.n_points, 3)),
              [("s", m.scalars.astype(np.float64)), ("part", parts + 1)],
              cell_arrays=[("partition_cells_1", lists[1])])
    j = last_json(run(probe, "load", path, "s", "part"))
    assert j["partition_cells"] == []


def probe_part(probe, tmp_path, cells, V, s, parts, backend, part_cells=None):
    files = []
    for name, a, dt in (("c.i64", cells, np.int64), ("s.f64", s, np.float64),
                        ("p.i32", parts, np.int32)):
        f = tmp_path / name
        np.ascontiguousarray(a, dt).tofile(f)
        files.append(f)
    args = ["part", files[0], V, files[1], files[2], backend]
    if part_cells is not None:
        f = tmp_path / "pc.i32"
        np.ascontiguousarray(part_cells, np.int32).tofile(f)
        args.append(f)
    return np.array(last_json(run(probe, *args))["classes"], np.uint32).reshape(-1, 3)


def partition_cases():
    rng = np.random.default_rng(8)
    soup = np.array([rng.choice(50, 4, replace=False) for _ in range(160)])
    s_soup = rng.integers(0, 5, 50).astype(np.float64)
    k = t.kuhn(5, seed=6, scalar="gaussians")
    hub = t.kuhn(4, seed=2, split_fraction=0.3, n_hubs=3, hub_splits=20, scalar="sines")
    nf = np.array([[0, 1, 2, 3], [0, 4, 5, 6], [1, 4, 7, 8]])
    return [
        ("soup4", soup, 50, s_soup, np.arange(50) % 4, None),
        ("soup_lists", soup, 50, s_soup, np.arange(50) % 3,
         rng.integers(0, 2, (3, len(soup)))),
        ("kuhn5", k.cells, k.n_points, k.scalars.astype(np.float64),
         rng.integers(0, 6, k.n_points), None),
        ("hubs", hub.cells, hub.n_points, hub.scalars.astype(np.float64),
         rng.integers(0, 3, hub.n_points), None),
        ("non_flag", nf, 10, np.array([10, 1, 2, 3, 4, 5, 6, 20, 20, 0], np.float64),
         np.array([0, 1, 1, 1, 1, 1, 1, 1, 1, 1]), None),
    ]


@pytest.mark.parametrize("case", partition_cases(), ids=lambda c: c[0])
def test_cpu_partitioned_critpoints_equal_oracle(probe, tmp_path, case):
    _, cells, V, s, parts, pc = case
    got = probe_part(probe, tmp_path, cells, V, s, parts, "cpu", pc)
    assert np.array_equal(got, O.critpoints_partitioned(cells, V, s, parts, pc))


def test_partitioned_classes_rejects_bad_ids_before_any_device_call():
    # argument checks of tetra_partitioned_classes run on the host first
    with pytest.raises(ValueError, match="negative partition id at point 2"):
        t.partitioned_classes(np.array([[0, 1, 2, 3]], np.int32), 4, np.zeros(4),
                              np.array([0, 0, -1, 0]))
    with pytest.raises(ValueError, match="one entry per point"):
        t.partitioned_classes(np.array([[0, 1, 2, 3]], np.int32), 4, np.zeros(3), np.zeros(4))
