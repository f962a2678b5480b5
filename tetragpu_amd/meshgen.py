"""Synthetic tetrahedral meshes (wrapper of include/tetra_meshgen.h).

Kuhn n^3 cube grids, 6 tets per cube, with seeded vertex and tet
permutations, +-0.2 cell jitter, optional stellar 1->4 splits and scalar
fields; the same seed gives the same mesh everywhere.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _native

PERMUTE_VERTICES = 0x1
PERMUTE_CELLS = 0x2
JITTER = 0x4
ORDER_SHIFT = 4
ORDERS = {"sorted": 0, "path": 1, "shuffled": 2}
SCALARS = {"uniform": 0, "gaussians": 1, "sines": 2, "linear": 3, "constant": 4}


class _GenMesh(C.Structure):
    _fields_ = [
        ("n_points", C.c_int64),
        ("n_cells", C.c_int64),
        ("cells", C.POINTER(C.c_int32)),
        ("positions", C.POINTER(C.c_float)),
        ("scalars", C.POINTER(C.c_float)),
        ("max_vt", C.c_int32),
    ]


@dataclass
class TetMesh:
    cells: np.ndarray       # (T, 4) int32
    n_points: int
    positions: np.ndarray   # (V, 3) float32
    scalars: np.ndarray     # (V,) float32
    max_vt: int

    @property
    def n_cells(self) -> int:
        return int(self.cells.shape[0])


def kuhn(n: int, seed: int = 1234, *, permute_vertices: bool = True,
         permute_cells: bool = True, jitter: bool = True, order: str = "sorted",
         split_fraction: float = 0.0, n_hubs: int = 0, hub_splits: int = 0,
         scalar: str = "uniform") -> TetMesh:
    opts = ((PERMUTE_VERTICES if permute_vertices else 0)
            | (PERMUTE_CELLS if permute_cells else 0)
            | (JITTER if jitter else 0)
            | (ORDERS[order] << ORDER_SHIFT))
    lib = _native.host()
    ptr = C.c_void_p()
    rc = lib.tg_gen_kuhn(n, seed, opts, split_fraction, n_hubs, hub_splits, SCALARS[scalar],
                         C.byref(ptr))
    if rc != 0:
        raise ValueError(f"tg_gen_kuhn failed with status {rc}")
    try:
        g = C.cast(ptr, C.POINTER(_GenMesh)).contents
        T, V = g.n_cells, g.n_points
        cells = np.ctypeslib.as_array(g.cells, shape=(max(T, 1) * 4,))[: 4 * T].reshape(T, 4).copy()
        pos = np.ctypeslib.as_array(g.positions, shape=(V * 3,)).reshape(V, 3).copy()
        sc = np.ctypeslib.as_array(g.scalars, shape=(V,)).copy()
        return TetMesh(cells, int(V), pos, sc, int(g.max_vt))
    finally:
        lib.tg_gen_free(ptr)


def kuhn_counts(n: int) -> dict:
    """Closed forms of an unsplit Kuhn n^3 grid with ascending cells."""
    return {"V": (n + 1) ** 3, "T": 6 * n ** 3, "E": 7 * n ** 3 + 9 * n ** 2 + 3 * n,
            "F": 12 * n ** 3 + 6 * n ** 2}
