"""Batches of independent meshes through one launch sequence per wave
(tg_batch_* in include/tetra_c.h)."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _native
from ._native import check


class Batch:
    """meshes: list of (cells (T_i, 4) int32 local ids, n_points_i, scalars_i or None).
    All meshes need scalars, or none."""

    def __init__(self, meshes, *, max_wave_cells: int = 0, stream: int | None = None,
                 f64: bool = False):
        self._lib = _native.hip()
        self._h = C.c_void_p()
        n = len(meshes)
        self.n_meshes = n
        self.cell_offsets = np.zeros(n + 1, np.int64)
        self.point_offsets = np.zeros(n + 1, np.int64)
        for i, (cells, V, _) in enumerate(meshes):
            self.cell_offsets[i + 1] = self.cell_offsets[i] + len(cells)
            self.point_offsets[i + 1] = self.point_offsets[i] + V
        cells = np.ascontiguousarray(np.concatenate([np.asarray(c, np.int32).reshape(-1, 4)
                                                     for c, _, _ in meshes]), np.int32)
        has_s = meshes[0][2] is not None
        sc = None
        if has_s:
            sc = np.ascontiguousarray(np.concatenate([np.asarray(s) for _, _, s in meshes]),
                                      np.float64 if f64 else np.float32)
        check(self._lib.tg_batch_create(C.byref(self._h), n, self.cell_offsets.ctypes.data,
                                        cells.ctypes.data, self.point_offsets.ctypes.data,
                                        None if sc is None else sc.ctypes.data, int(f64),
                                        int(max_wave_cells), stream))

    def close(self):
        if self._h:
            self._lib.tg_batch_destroy(self._h)
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

    def counts(self) -> dict:
        w, c, p = C.c_int32(), C.c_int64(), C.c_int64()
        check(self._lib.tg_batch_counts(self._h, C.byref(w), C.byref(c), C.byref(p)))
        return {"waves": w.value, "cells": c.value, "points": p.value}

    def use_graph(self, on: bool = True) -> None:
        check(self._lib.tg_batch_use_graph(self._h, int(on)))

    def set_config(self, variant=0, lane_cap=32, lane_block=128, bv_log=7, tets_per_chunk=8192):
        check(self._lib.tg_batch_set_config(self._h, variant, lane_cap, lane_block, bv_log,
                                            tets_per_chunk))

    def vv_critpoints(self) -> None:
        check(self._lib.tg_batch_vv_critpoints(self._h))

    def sync(self) -> None:
        check(self._lib.tg_batch_sync(self._h))

    def classes(self) -> np.ndarray:
        out = np.empty((int(self.point_offsets[-1]), 3), np.uint32)
        check(self._lib.tg_batch_classes_copy(self._h, out.ctypes.data))
        return out

    def mesh_classes(self, i: int, all_classes: np.ndarray | None = None) -> np.ndarray:
        a = self.classes() if all_classes is None else all_classes
        return a[self.point_offsets[i]:self.point_offsets[i + 1]]

    def vv(self, i: int):
        nnz = C.c_int64()
        V = int(self.point_offsets[i + 1] - self.point_offsets[i])
        check(self._lib.tg_batch_vv_copy(self._h, i, None, None, C.byref(nnz)))
        off = np.empty(V + 1, np.int64)
        vals = np.empty(nnz.value, np.int32)
        check(self._lib.tg_batch_vv_copy(self._h, i, off.ctypes.data, vals.ctypes.data,
                                         C.byref(nnz)))
        return off, vals
