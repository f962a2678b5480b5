"""ctypes binding of the CPU oracle (oracle/oracle.c) -- test infrastructure.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use this
module, as the checker; the product never imports it.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "lib" / "liboracle.so"


class _Csr(C.Structure):
    _fields_ = [("rows", C.c_int64), ("nnz", C.c_int64), ("width", C.c_int32),
                ("off", C.POINTER(C.c_int64)), ("val", C.POINTER(C.c_int64))]


_lib = None


def build() -> None:
    r = subprocess.run(["make", "-s", "-C", str(HERE)], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("oracle build failed:\n" + r.stdout + r.stderr)


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIB.exists():
            build()
        L = C.CDLL(str(LIB))
        P, I64 = C.c_void_p, C.c_int64
        CP = C.POINTER(_Csr)
        sigs = {
            "orc_csr_free": (None, [CP]),
            "orc_make_TE_and_VE": (I64, [P, I64, I64, P, CP]),
            "orc_make_TF_and_VF": (I64, [P, I64, I64, P, CP]),
            "orc_make_EV": (None, [CP, I64, P]),
            "orc_make_ET": (None, [P, I64, I64, CP]),
            "orc_make_FV": (None, [CP, I64, P]),
            "orc_make_VV": (None, [P, I64, I64, CP]),
            "orc_make_VT": (None, [P, I64, I64, CP]),
            "orc_make_FE": (C.c_int, [P, I64, CP, P]),
            "orc_make_FT": (C.c_int, [P, I64, I64, P]),
            "orc_make_EF": (None, [P, I64, I64, CP]),
            "orc_make_TT": (None, [P, I64, I64, CP]),
            "orc_make_FF": (None, [P, I64, P, I64, CP]),
            "orc_make_EE": (None, [P, I64, I64, CP]),
            "orc_critpoints": (None, [CP, P, I64, P]),
            "orc_label6": (None, [P, I64, P]),
            "orc_approx_max_vv": (C.c_int32, [P, I64, I64]),
            "orc_vv_cp_omp": (C.c_double, [P, I64, I64, P, C.c_int, CP, P]),
        }
        for n, (res, args) in sigs.items():
            f = getattr(L, n)
            f.restype, f.argtypes = res, args
        _lib = L
    return _lib


class Csr:
    """CSR relation: off int64 [rows+1], val int64 [nnz, width]."""

    def __init__(self, off: np.ndarray, val: np.ndarray):
        self.off, self.val = off, val

    @property
    def rows(self) -> int:
        return len(self.off) - 1

    def row(self, i: int) -> np.ndarray:
        return self.val[self.off[i]:self.off[i + 1]]

    def sorted_rows(self) -> "Csr":
        v = self.val.copy()
        for i in range(self.rows):
            a, b = self.off[i], self.off[i + 1]
            if v.shape[1] == 1:
                v[a:b, 0] = np.sort(v[a:b, 0])
        return Csr(self.off.copy(), v)

    def to_c(self) -> _Csr:
        self._keep = (np.ascontiguousarray(self.off, np.int64),
                      np.ascontiguousarray(self.val, np.int64))
        c = _Csr()
        c.rows = self.rows
        c.nnz = self.val.shape[0]
        c.width = self.val.shape[1] if self.val.ndim == 2 else 1
        c.off = self._keep[0].ctypes.data_as(C.POINTER(C.c_int64))
        c.val = self._keep[1].ctypes.data_as(C.POINTER(C.c_int64))
        return c


_libc = C.CDLL(None)
_libc.free.argtypes = [C.c_void_p]
_libc.free.restype = None


def _adopt(ptr, n: int) -> np.ndarray:
    """numpy view of n int64 words malloc'ed by the oracle, freed with the
    array (no copy: TT of a 12.6M-cell mesh holds about 7 GB)."""
    import weakref

    a = np.ctypeslib.as_array(ptr, shape=(max(n, 1),))[:n]
    weakref.finalize(a.base if a.base is not None else a, _libc.free,
                     C.cast(ptr, C.c_void_p).value)
    return a


def _take(c: _Csr) -> Csr:
    rows, nnz, w = c.rows, c.nnz, c.width
    off = _adopt(c.off, rows + 1)
    val = _adopt(c.val, nnz * w).reshape(nnz, w)
    c.off = c.val = None
    return Csr(off, val)


def _i64(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.int64)


def edges(cells, n_points):
    """(E, TE (T,6), VE csr {hi,id}) by make_TE_and_VE."""
    cells = _i64(cells).reshape(-1, 4)
    T = cells.shape[0]
    te = np.empty((T, 6), np.int64)
    c = _Csr()
    E = lib().orc_make_TE_and_VE(cells.ctypes.data, T, n_points, te.ctypes.data, C.byref(c))
    if E < 0:
        raise ValueError("invalid cells")
    return int(E), te, _take(c)


def faces(cells, n_points):
    """(F, TF (T,4), VF csr {mid,hi,id}) by make_TF_and_VF."""
    cells = _i64(cells).reshape(-1, 4)
    T = cells.shape[0]
    tf = np.empty((T, 4), np.int64)
    c = _Csr()
    F = lib().orc_make_TF_and_VF(cells.ctypes.data, T, n_points, tf.ctypes.data, C.byref(c))
    if F < 0:
        raise ValueError("invalid cells")
    return int(F), tf, _take(c)


def ev(VE: Csr, E: int) -> np.ndarray:
    out = np.empty((E, 2), np.int64)
    c = VE.to_c()
    lib().orc_make_EV(C.byref(c), E, out.ctypes.data)
    return out


def et(TE, E: int) -> Csr:
    TE = _i64(TE)
    c = _Csr()
    lib().orc_make_ET(TE.ctypes.data, TE.shape[0], E, C.byref(c))
    return _take(c)


def fv(VF: Csr, F: int) -> np.ndarray:
    """FV rows as FaceData fields (middleVert, highVert, id := low vertex)."""
    out = np.empty((F, 3), np.int64)
    c = VF.to_c()
    lib().orc_make_FV(C.byref(c), F, out.ctypes.data)
    return out


def vv(cells, n_points) -> Csr:
    cells = _i64(cells).reshape(-1, 4)
    c = _Csr()
    lib().orc_make_VV(cells.ctypes.data, cells.shape[0], n_points, C.byref(c))
    return _take(c)


def vt(cells, n_points) -> Csr:
    cells = _i64(cells).reshape(-1, 4)
    c = _Csr()
    lib().orc_make_VT(cells.ctypes.data, cells.shape[0], n_points, C.byref(c))
    return _take(c)


def fe(FV, VE: Csr) -> np.ndarray:
    """[F][3] of (highVert, edge id)."""
    FV = _i64(FV)
    out = np.empty((FV.shape[0], 3, 2), np.int64)
    c = VE.to_c()
    if lib().orc_make_FE(FV.ctypes.data, FV.shape[0], C.byref(c), out.ctypes.data) != 0:
        raise ValueError("face edge missing from VE")
    return out


def ft(TF, F: int) -> np.ndarray:
    TF = _i64(TF)
    out = np.empty((F, 2), np.int64)
    if lib().orc_make_FT(TF.ctypes.data, TF.shape[0], F, out.ctypes.data) != 0:
        raise ValueError("a face is held by more than two tets")
    return out


def ef(FE, E: int) -> Csr:
    FE = _i64(FE)
    c = _Csr()
    lib().orc_make_EF(FE.ctypes.data, FE.shape[0], E, C.byref(c))
    return _take(c)


def tt(cells, n_points) -> Csr:
    cells = _i64(cells).reshape(-1, 4)
    c = _Csr()
    lib().orc_make_TT(cells.ctypes.data, cells.shape[0], n_points, C.byref(c))
    return _take(c)


def ff(TF, FT) -> Csr:
    TF, FT = _i64(TF), _i64(FT)
    c = _Csr()
    lib().orc_make_FF(TF.ctypes.data, TF.shape[0], FT.ctypes.data, FT.shape[0], C.byref(c))
    return _take(c)


def ee(EV, n_points) -> Csr:
    EV = _i64(EV)
    c = _Csr()
    lib().orc_make_EE(EV.ctypes.data, EV.shape[0], n_points, C.byref(c))
    return _take(c)


def critpoints(VV: Csr, scalars, n_points) -> np.ndarray:
    sc = np.ascontiguousarray(scalars, dtype=np.float64)
    out = np.empty((n_points, 3), np.uint32)
    c = VV.to_c()
    lib().orc_critpoints(C.byref(c), sc.ctypes.data, n_points, out.ctypes.data)
    return out


def critpoints_partitioned(cells, n_points, scalars, parts, part_cells=None) -> np.ndarray:
    """classes (V, 3) of the reference's partition loop: partition p is
    classified on its cells -- part_cells[p] (a 0/1 row per cell) when given
    (src/alg/critPoints.cu:731-748), else every cell with a vertex in p
    (:814-838) -- with vertices renumbered in ascending global order
    (:846-905); only p's own vertices keep their classes (:940-943) and
    results merge by ownership (:1462-1491).  Vertices in no cell of their
    partition stay (0, 0, 0)."""
    cells = _i64(cells).reshape(-1, 4)
    parts = np.ascontiguousarray(parts, dtype=np.int64)
    sc = np.ascontiguousarray(scalars, dtype=np.float64)
    out = np.zeros((n_points, 3), np.uint32)
    n_parts = int(parts.max()) + 1 if parts.size else 1
    for p in range(n_parts):
        if part_cells is not None:
            mask = np.asarray(part_cells[p]) == 1
        else:
            mask = (parts[cells] == p).any(axis=1) if cells.size else np.zeros(0, bool)
        sub = cells[mask]
        if sub.size == 0:
            continue
        verts = np.unique(sub)
        local = np.searchsorted(verts, sub)
        cls = critpoints(vv(local, verts.size), sc[verts], verts.size)
        own = parts[verts] == p
        out[verts[own]] = cls[own]
    return out


def label6(classes) -> np.ndarray:
    cl = np.ascontiguousarray(classes, dtype=np.uint32)
    out = np.empty(cl.shape[0], np.uint8)
    lib().orc_label6(cl.ctypes.data, cl.shape[0], out.ctypes.data)
    return out


def approx_max_vv(cells, n_points) -> int:
    cells = _i64(cells).reshape(-1, 4)
    return int(lib().orc_approx_max_vv(cells.ctypes.data, cells.shape[0], n_points))


def vv_cp_omp(cells, n_points, scalars, threads: int = 0):
    """(VV csr with ascending rows, classes (V,3), seconds) -- OpenMP form."""
    cells = np.ascontiguousarray(cells, dtype=np.int32).reshape(-1, 4)
    sc = np.ascontiguousarray(scalars, dtype=np.float32)
    out = np.empty((n_points, 3), np.uint32)
    c = _Csr()
    secs = lib().orc_vv_cp_omp(cells.ctypes.data, cells.shape[0], n_points, sc.ctypes.data,
                               threads, C.byref(c), out.ctypes.data)
    return _take(c), out, float(secs)
