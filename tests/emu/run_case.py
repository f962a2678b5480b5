"""Runs one GPU test case on the CPU kernel emulator.

Started by tests/test_emu_kernels.py as a child process with the
AddressSanitizer runtime preloaded; the package's ctypes layer is pointed at
tests/emu/lib/libtetra_hip_emu.so for this process only.  Prints "ok <case>"
and exits 0 when the case's results equal the oracle's; an out-of-bounds
access aborts the process with ASan's report.

Usage: python tests/emu/run_case.py <case> [args...]
"""
from __future__ import annotations

import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
ROOT = HERE.parent.parent
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "tests"))

import numpy as np  # noqa: E402

import tetragpu_amd as t  # noqa: E402
from tetragpu_amd import _native  # noqa: E402


def use_emulator() -> None:
    import os

    path = os.environ.get("TETRA_EMU_LIB", str(HERE / "lib" / "libtetra_hip_emu.so"))
    lib = _native.C.CDLL(path, mode=_native.C.RTLD_GLOBAL)
    for fn, (res, args) in _native._HIP_SIGS.items():
        f = getattr(lib, fn, None)  # an older build may lack newer entry points
        if f is not None:
            f.restype, f.argtypes = res, args
    _native._hip = lib


def case_after_hot_path(seed: int = 8, n: int = 5, variant: int = 1) -> None:
    """TE, then the hot path, then ET/TT/VF on one mesh."""
    from test_gpu_relations import oracle_all, same_csr, same_fixed

    m = t.kuhn(n, seed=seed)
    ref = oracle_all(m.cells, m.n_points)
    with t.Mesh(m.cells, m.n_points, m.scalars) as d:
        d.set_config(variant=variant)
        d.build(t.flags("TE"))
        d.vv_critpoints()
        d.build(t.flags("ET", "TT", "VF"))
        same_csr(d.relation("ET"), ref["ET"], "ET")
        same_csr(d.relation("TT"), ref["TT"], "TT")
        same_csr(d.relation("VF"), ref["VF"], "VF")
        same_fixed(d.relation("TE"), ref["TE"], "TE")


def case_all_relations(seed: int = 3, n: int = 4, variant: int = 0) -> None:
    """Every relation flag in one build, against the oracle."""
    from test_gpu_relations import check_all

    m = t.kuhn(n, seed=seed)
    check_all(t, m.cells, m.n_points, variant)


def case_vv_cp(seed: int = 5, n: int = 6, variant: int = 2) -> None:
    """VV + classes in one kernel form, against the oracle."""
    m = t.kuhn(n, seed=seed)
    same_vv_cp(m.cells, m.n_points, m.scalars, variant)


def same_vv_cp(cells, V, scalars, variant: int) -> None:
    from oracle import binding as O

    with t.Mesh(cells, V, scalars) as d:
        d.set_config(variant=variant)
        d.vv_critpoints()
        off, vals = d.relation("VV")
        cls = d.classes()
    VV = O.vv(cells, V)
    S = VV.sorted_rows()
    assert np.array_equal(off, S.off) and np.array_equal(vals[:, 0], S.val[:, 0]), "VV"
    ref = O.critpoints(VV, np.asarray(scalars, np.float64), V)
    assert np.array_equal(cls, ref), "classes"


def case_window(seed: int = 1, V: int = 100, T: int = 760, window: int = 16,
                variant: int = 2) -> None:
    """A soup whose cells take four vertices from a window of `window`
    consecutive ids: rows stay short (at most 2 * window - 2 neighbours)
    while vertices carry many records, so buckets fill up with records of
    few distinct keys (and, for wide windows, rows go to the heavy kernel)."""
    rng = np.random.default_rng(seed)
    w0 = rng.integers(0, V - window + 1, T)
    cells = np.array([w + rng.choice(window, 4, replace=False) for w in w0], np.int32)
    scalars = rng.random(V).astype(np.float32)
    same_vv_cp(cells, V, scalars, variant)


def case_hubs(seed: int = 11, hub_splits: int = 40) -> None:
    """High-valence hubs: rows beyond the LDS columns take the global paths."""
    from test_gpu_relations import check_all

    m = t.kuhn(4, seed=seed, split_fraction=0.2, n_hubs=3, hub_splits=hub_splits)
    check_all(t, m.cells, m.n_points)


def case_soup(seed: int = 2, V: int = 60, T: int = 200) -> None:
    """A random cell soup (non-manifold, faces over more than two cells)."""
    from test_gpu_relations import check_all

    rng = np.random.default_rng(seed)
    cells = np.array([rng.choice(V, 4, replace=False) for _ in range(T)], np.int32)
    check_all(t, cells, V, expect_ft=False)


def case_routes(seed: int = 4, route: int = 2, bv: int = 7) -> None:
    """VV + classes and VT through one routing of the bucketed form."""
    from oracle import binding as O

    m = t.kuhn(7, seed=seed)
    with t.Mesh(m.cells, m.n_points, m.scalars) as d:
        d.set_config(variant=2)
        d.set_bucket(bv, 4096)
        d.set_routing(route)
        d.vv_critpoints()
        d.build(t.flags("VT"))
        cls = d.classes()
        vt = d.relation("VT")
    VV = O.vv(m.cells, m.n_points)
    assert np.array_equal(cls, O.critpoints(VV, m.scalars.astype(np.float64), m.n_points))
    ref = O.vt(m.cells, m.n_points)
    assert np.array_equal(vt[0], ref.off) and np.array_equal(vt[1], ref.val), "VT"


def case_noalloc(seed: int = 6, n1: int = 7, n2: int = 4) -> None:
    """The no-allocation entry point (host arrays stand in for device memory
    here) against the mesh path, two calls in a row, the second smaller."""
    C = _native.C
    lib = _native.hip()
    for n in (n1, n2):
        m = t.kuhn(n, seed=seed + n)
        with t.Mesh(m.cells, m.n_points, m.scalars) as d:
            d.vv_critpoints()
            off_ref, vals_ref = d.relation("VV")
            cls_ref = d.classes()
        T, V = m.n_cells, m.n_points
        cells = np.ascontiguousarray(m.cells, np.int32)
        sc = np.ascontiguousarray(m.scalars, np.float32)
        cap = min(12 * T, V * (V - 1))
        off = np.full(V + 1, -7, np.int32)
        cols = np.full(cap, -7, np.int32)
        cls = np.full((V, 3), 7, np.uint32)
        st = np.full(1, -7, np.int32)
        _native.check(lib.tg_vv_critpoints_noalloc(cells.ctypes.data, T, V, sc.ctypes.data, 0,
                                                   off.ctypes.data, cols.ctypes.data, cap,
                                                   cls.ctypes.data, st.ctypes.data, None))
        assert st[0] == 0
        assert np.array_equal(off, off_ref), "VV offsets"
        assert np.array_equal(cols[: off[-1]], vals_ref[:, 0]), "VV columns"
        assert np.array_equal(cls, cls_ref), "classes"


def split_budget(T: int, V: int, frac: float) -> int:
    """A device budget that holds the resident arrays of a split run plus
    `frac` of the whole mesh's footprint."""
    resident = 16 * T + 8 * V + 12 * V + 4 * (T + 1) + 4 * (V + 1) + V + (16 << 10) + 2048
    return resident + int(frac * t.vv_footprint(T, V))


def case_split(seed: int = 3, n: int = 4) -> None:
    """Critical points in parts, localised on the device: vertex-id ranges
    forced by a small budget equal the whole mesh (a stellar-split mesh with
    an isolated vertex); caller partitions equal the oracle's partition loop
    (explicit cell rows, a non-flag link); budgets that cannot hold the
    resident arrays or a partition fail with TG_ERR_ALLOC."""
    from oracle import binding as O
    from test_dropin_cpu import partition_cases

    m = t.kuhn(n, seed=seed, permute_vertices=False, permute_cells=False, split_fraction=0.3,
               scalar="sines")
    cells, V = m.cells, m.n_points + 1          # one vertex in no cell
    s = np.append(m.scalars.astype(np.float64), 0.5)
    whole = O.critpoints(O.vv(cells, V), s, V)
    assert whole[-1].tolist() == [0, 0, 4]
    got, st = t.split_classes(cells, V, s, budget_bytes=split_budget(len(cells), V, 0.6))
    assert st["parts"] >= 2, st
    assert np.array_equal(got, whole), ("split", st)
    assert st["cells_total"] > len(cells)       # halo cells are counted again
    for name, cells_p, Vp, sp, parts, pc in partition_cases():
        if name in ("soup_lists", "non_flag"):
            got, st = t.split_classes(cells_p, Vp, sp, parts=parts, part_cells=pc)
            assert np.array_equal(got, O.critpoints_partitioned(cells_p, Vp, sp, parts, pc)), name
    # an id outside [0, V) is rejected before any kernel indexes with it
    bad = cells.copy()
    bad[len(bad) // 2, 1] = V
    for kw in (dict(budget_bytes=split_budget(len(cells), V, 0.6)),
               dict(parts=np.zeros(V, np.int32))):
        try:
            t.split_classes(bad, V, s, **kw)
        except _native.NativeError as e:
            assert e.status == -4 and f"cell {len(bad) // 2}" in str(e), e
        else:
            raise AssertionError("expected TG_ERR_MESH")
    for kw in (dict(budget_bytes=4096),
               dict(parts=np.zeros(V, np.int32), budget_bytes=split_budget(len(cells), V, 0.2))):
        try:
            t.split_classes(cells, V, s, **kw)
        except _native.NativeError as e:
            assert e.status == -3, e
        else:
            raise AssertionError(f"expected TG_ERR_ALLOC for {list(kw)}")


CASES = {"split": case_split, "noalloc": case_noalloc, "after_hot_path": case_after_hot_path, "all_relations": case_all_relations,
         "vv_cp": case_vv_cp, "window": case_window, "hubs": case_hubs, "soup": case_soup, "routes": case_routes}


def main() -> int:
    import os

    name = sys.argv[1]
    if name.endswith("_y"):  # preempt work-items at every third atomic (lost CAS races)
        os.environ["EMU_YIELD"] = "3"
        name = name[:-2]
    elif name.endswith("_t"):  # workgroups of a launch on 4 host threads at once
        os.environ["EMU_THREADS"] = "4"
        name = name[:-2]
    elif name.endswith("_r"):  # workgroups from the last one down
        os.environ["EMU_ORDER"] = "reverse"
        name = name[:-2]
    args = [int(a) for a in sys.argv[2:]]
    use_emulator()
    t.init(0)
    CASES[name](*args)
    print(f"ok {sys.argv[1]} {' '.join(sys.argv[2:])}", flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
