"""Python view of the C ABI (include/tetra_c.h): a device-resident mesh, its
relations and critical points.

Relation ids are the bit positions of the reference's runtime_arguments
flags (include/datatypes.h.in:158-178); `flags(...)` builds a mask from
names.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _native
from ._native import check

REL = {"TE": 1, "EV": 2, "ET": 3, "TF": 4, "FV": 5, "FE": 6, "FT": 7, "EF": 8,
       "VT": 9, "TT": 10, "FF": 11, "EE": 12, "VV": 13, "VE": 14, "VF": 15}
ALL_FLAGS = 0x3FFE

CLASS_NAMES = {1: "minimum", 2: "maximum", 3: "regular", 4: "saddle"}
LABEL6_NAMES = {0: "minimum", 1: "1-saddle", 2: "2-saddle", 3: "maximum",
                4: "multi-saddle", 5: "regular"}


def flags(*names: str) -> int:
    m = 0
    for n in names:
        m |= 1 << REL[n]
    return m


def init(device: int = 0) -> None:
    check(_native.hip().tg_init(device))


def partitioned_classes(cells, n_points: int, scalars, parts):
    """classes (V, 3) from the reference's partition loop on the device, one
    partition's cells at a time (src/alg/critPoints.cu:654-1003, merged as at
    :1438-1491; libtetra_host.so, tetra_partitioned_classes).  Partition p
    is classified on the cells that hold one of its vertices; only its own
    vertices keep their classes.  Returns (classes, stats)."""
    cells = np.ascontiguousarray(cells, dtype=np.int32).reshape(-1, 4)
    sc = np.ascontiguousarray(scalars, dtype=np.float64)
    pa = np.ascontiguousarray(parts, dtype=np.int32)
    if sc.shape[0] != n_points or pa.shape[0] != n_points:
        raise ValueError("scalars and parts need one entry per point")
    out = np.zeros((n_points, 3), np.uint32)
    st = np.zeros(7, np.float64)
    rc = _native.host().tetra_partitioned_classes(cells.ctypes.data, cells.shape[0], n_points,
                                                  sc.ctypes.data, pa.ctypes.data,
                                                  out.ctypes.data, st.ctypes.data)
    if rc != 0:
        msg = _native.host().tetra_partition_error().decode(errors="replace")
        if rc == -1:
            raise ValueError(msg)
        raise _native.NativeError(rc, msg)
    keys = ("partitions", "total_s", "localize_s", "classify_s", "max_cells", "max_points",
            "cells_total")
    stats = {k: (int(v) if k in ("partitions", "max_cells", "max_points", "cells_total")
                 else float(v)) for k, v in zip(keys, st)}
    return out, stats


class _SplitStats(C.Structure):
    _fields_ = [("parts", C.c_int64), ("max_cells", C.c_int64), ("max_points", C.c_int64),
                ("cells_total", C.c_int64), ("footprint_bytes", C.c_int64),
                ("budget_bytes", C.c_int64), ("total_s", C.c_double),
                ("localize_s", C.c_double), ("classify_s", C.c_double)]


def vv_footprint(n_cells: int, n_points: int) -> int:
    """Device bytes the VV + critical-point path allocates for a mesh of this
    size (tg_vv_footprint)."""
    b = C.c_int64(0)
    _native.check(_native.hip().tg_vv_footprint(int(n_cells), int(n_points), C.byref(b)))
    return int(b.value)


def split_classes(cells, n_points: int, scalars, parts=None, part_cells=None,
                  budget_bytes: int = 0):
    """classes (V, 3) computed in parts, each localised on the device
    (tg_critpoints_split).  parts=None: the whole mesh when its footprint
    fits budget_bytes (0: 90 % of free device memory), else vertex-id ranges
    with their neighbours' cells, identical to the whole mesh.  parts (one id
    per point): the reference's partition loop, optionally with explicit cell
    rows part_cells (n_parts, T).  Returns (classes, stats)."""
    cells = np.ascontiguousarray(cells, dtype=np.int32).reshape(-1, 4)
    sc = np.ascontiguousarray(scalars)
    if sc.dtype not in (np.float32, np.float64):
        sc = sc.astype(np.float64)
    if sc.shape[0] != n_points:
        raise ValueError("scalars need one entry per point")
    pa_ptr, n_parts, rows_ptr = None, 0, None
    if parts is not None:
        pa = np.ascontiguousarray(parts, dtype=np.int32)
        if pa.shape[0] != n_points:
            raise ValueError("parts need one entry per point")
        n_parts = int(pa.max()) + 1 if n_points else 1
        pa_ptr = pa.ctypes.data
        if part_cells is not None:
            rows = np.ascontiguousarray(part_cells, dtype=np.int32)
            if rows.shape != (n_parts, cells.shape[0]):
                raise ValueError("part_cells must be (n_parts, n_cells)")
            rows_ptr = rows.ctypes.data
    elif part_cells is not None:
        raise ValueError("part_cells needs parts")
    out = np.zeros((n_points, 3), np.uint32)
    st = _SplitStats()
    _native.check(_native.hip().tg_critpoints_split(
        cells.ctypes.data, cells.shape[0], int(n_points), sc.ctypes.data,
        1 if sc.dtype == np.float64 else 0, pa_ptr, n_parts, rows_ptr, int(budget_bytes),
        out.ctypes.data, C.byref(st)))
    stats = {k: getattr(st, k) for k, _ in _SplitStats._fields_}
    return out, stats


def device_count() -> int:
    n = C.c_int(0)
    _native.hip().tg_device_count(C.byref(n))
    return int(n.value)


class Mesh:
    """Device-resident tetrahedral mesh (tg_mesh handle).

    cells: (T, 4) int32 host array, or a device pointer with n_cells given and
    on_device=True.  scalars: (V,) float32 or float64, optional.
    stream: hipStream_t as an int (e.g. torch.cuda.current_stream().cuda_stream)
    or None for a stream owned by the mesh.
    """

    def __init__(self, cells, n_points: int, scalars=None, *, stream: int | None = None,
                 on_device: bool = False, n_cells: int | None = None,
                 scalars_f64: bool | None = None):
        self._lib = _native.hip()
        self._h = C.c_void_p()
        self.n_points = int(n_points)
        if on_device:
            cells_ptr, T = int(cells), int(n_cells)
            sc_ptr = int(scalars) if scalars is not None else None
            f64 = bool(scalars_f64)
        else:
            cells = np.ascontiguousarray(cells, dtype=np.int32).reshape(-1, 4)
            self._cells_keep = cells
            cells_ptr, T = cells.ctypes.data, cells.shape[0]
            sc_ptr, f64 = None, False
            if scalars is not None:
                sc = np.asarray(scalars)
                f64 = sc.dtype == np.float64
                sc = np.ascontiguousarray(sc, dtype=np.float64 if f64 else np.float32)
                self._sc_keep = sc
                sc_ptr = sc.ctypes.data
        self.n_cells = T
        check(self._lib.tg_mesh_create(C.byref(self._h), cells_ptr, T, self.n_points,
                                       sc_ptr, int(f64), int(on_device), stream))

    # -- lifecycle -----------------------------------------------------------
    def close(self) -> None:
        if self._h:
            self._lib.tg_mesh_destroy(self._h)
            self._h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self) -> C.c_void_p:
        return self._h

    @property
    def stream(self) -> int:
        s = C.c_void_p()
        check(self._lib.tg_mesh_stream(self._h, C.byref(s)))
        return int(s.value or 0)

    # -- work ------------------------------------------------------------------
    def set_scalars(self, scalars) -> None:
        sc = np.asarray(scalars)
        f64 = sc.dtype == np.float64
        sc = np.ascontiguousarray(sc, dtype=np.float64 if f64 else np.float32)
        self._sc_keep = sc
        check(self._lib.tg_mesh_set_scalars(self._h, sc.ctypes.data, int(f64), 0))

    def set_config(self, lane_cap: int = 32, lane_block: int = 128, variant: int = 0) -> None:
        check(self._lib.tg_mesh_set_config(self._h, variant, lane_cap, lane_block))

    def set_bucket(self, bv_log: int = 7, tets_per_chunk: int = 8192) -> None:
        """Bucketed form: 2^bv_log vertices per bucket, tets per routing chunk."""
        check(self._lib.tg_mesh_set_bucket(self._h, bv_log, tets_per_chunk))

    def set_routing(self, route: int = 0) -> None:
        """Bucketed form's record routing: 0 auto, 1 one level, 2 two levels."""
        check(self._lib.tg_mesh_set_routing(self._h, route))

    def routing(self) -> int:
        r = C.c_int32()
        check(self._lib.tg_mesh_routing(self._h, C.byref(r)))
        return r.value

    def build(self, mask: int) -> None:
        check(self._lib.tg_build(self._h, mask))

    def critpoints(self) -> None:
        check(self._lib.tg_critpoints(self._h))

    def vv_critpoints(self) -> None:
        check(self._lib.tg_vv_critpoints(self._h))

    def sync(self) -> None:
        check(self._lib.tg_sync(self._h))

    # -- results ---------------------------------------------------------------
    def counts(self) -> dict:
        v, t, e, f = (C.c_int64() for _ in range(4))
        check(self._lib.tg_mesh_counts(self._h, C.byref(v), C.byref(t), C.byref(e), C.byref(f)))
        return {"V": v.value, "T": t.value, "E": e.value, "F": f.value}

    def shape(self, name: str) -> dict:
        rows, nnz, width, fixed = C.c_int64(), C.c_int64(), C.c_int32(), C.c_int32()
        check(self._lib.tg_relation_shape(self._h, REL[name], C.byref(rows), C.byref(nnz),
                                          C.byref(width), C.byref(fixed)))
        return {"rows": rows.value, "nnz": nnz.value, "width": width.value,
                "fixed": bool(fixed.value)}

    def relation(self, name: str):
        """(offsets int64 [rows+1] or None, values int32 [nnz, width])."""
        rel = REL[name]
        rows, nnz, width, fixed = C.c_int64(), C.c_int64(), C.c_int32(), C.c_int32()
        check(self._lib.tg_relation_shape(self._h, rel, C.byref(rows), C.byref(nnz),
                                          C.byref(width), C.byref(fixed)))
        vals = np.empty((nnz.value, width.value), dtype=np.int32)
        offs = None if fixed.value else np.empty(rows.value + 1, dtype=np.int64)
        check(self._lib.tg_relation_copy(self._h, rel,
                                         None if offs is None else offs.ctypes.data,
                                         vals.ctypes.data))
        return offs, vals

    def relation_device(self, name: str):
        off, ob, val = C.c_void_p(), C.c_int32(), C.c_void_p()
        check(self._lib.tg_relation_device(self._h, REL[name], C.byref(off), C.byref(ob),
                                           C.byref(val)))
        return off.value, ob.value, val.value

    def classes(self) -> np.ndarray:
        out = np.empty((self.n_points, 3), dtype=np.uint32)
        check(self._lib.tg_classes_copy(self._h, out.ctypes.data))
        return out

    def labels6(self) -> np.ndarray:
        out = np.empty(self.n_points, dtype=np.uint8)
        check(self._lib.tg_labels6_copy(self._h, out.ctypes.data))
        return out

    def profile(self, on: bool = True) -> None:
        check(self._lib.tg_profile_enable(self._h, int(on)))

    def stage_times(self) -> list[tuple[str, float]]:
        n = C.c_int32()
        ms = (C.c_float * 32)()
        names = C.create_string_buffer(32 * 32)
        check(self._lib.tg_profile_read(self._h, 32, C.byref(n), ms, names))
        raw = names.raw
        return [(raw[32 * i:32 * i + 32].split(b"\0")[0].decode(), float(ms[i]))
                for i in range(min(n.value, 32))]

    def approx_max_vv(self) -> int:
        r = C.c_int32()
        check(self._lib.tg_approx_max_vv(self._h, C.byref(r)))
        return int(r.value)
