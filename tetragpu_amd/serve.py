"""Request loop with continuous batching (tg_server_* in include/tetra/serve.h)."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _native
from .mesh import REL

VT = 1 << REL["VT"]
TT = 1 << REL["TT"]
COPY = 1 << 31  # TG_SERVE_COPY: follow-up relations also copied to host memory


class Config(C.Structure):
    _fields_ = [("slots", C.c_int32), ("hbm_budget_bytes", C.c_int64),
                ("fuse_max_cells", C.c_int64), ("fuse_max_meshes", C.c_int32)]


class Result(C.Structure):
    _fields_ = [("status", C.c_int32), ("n_points", C.c_int64), ("n_cells", C.c_int64),
                ("classes", C.POINTER(C.c_uint32)),
                ("vt_off", C.POINTER(C.c_int64)), ("vt_val", C.POINTER(C.c_int32)),
                ("vt_nnz", C.c_int64),
                ("tt_off", C.POINTER(C.c_int64)), ("tt_val", C.POINTER(C.c_int32)),
                ("tt_nnz", C.c_int64),
                ("queue_ms", C.c_double), ("run_ms", C.c_double), ("total_ms", C.c_double),
                ("group_size", C.c_int32), ("upload_ms", C.c_double), ("compute_ms", C.c_double),
                ("followup_ms", C.c_double), ("copy_ms", C.c_double)]


_SIGS = {
    "tg_server_create": (C.c_int, [C.POINTER(C.c_void_p), C.POINTER(Config)]),
    "tg_server_submit": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int64, C.c_int64, C.c_void_p,
                                   C.c_uint32, C.POINTER(C.c_uint64)]),
    "tg_server_wait": (C.c_int, [C.c_void_p, C.c_uint64, C.POINTER(Result)]),
    "tg_server_followup": (C.c_int, [C.c_void_p, C.c_uint64, C.c_uint32, C.POINTER(Result)]),
    "tg_server_release": (C.c_int, [C.c_void_p, C.c_uint64]),
    "tg_server_stats": (C.c_int, [C.c_void_p, C.POINTER(C.c_int64), C.POINTER(C.c_int64),
                                  C.POINTER(C.c_int64)]),
    "tg_server_destroy": (C.c_int, [C.c_void_p]),
}


def _lib():
    lib = _native.host()
    if not getattr(lib, "_serve_bound", False):
        for fn, (res, args) in _SIGS.items():
            f = getattr(lib, fn)
            f.restype, f.argtypes = res, args
        lib._serve_bound = True
    return lib


def _arr(ptr, n, dtype):
    return np.ctypeslib.as_array(ptr, shape=(n,)).astype(dtype, copy=True) if n > 0 \
        else np.zeros(0, dtype)


def result_dict(r: Result) -> dict:
    V, T = r.n_points, r.n_cells
    out = {"status": r.status, "classes": _arr(r.classes, 3 * V, np.uint32).reshape(V, 3),
           "queue_ms": r.queue_ms, "run_ms": r.run_ms, "total_ms": r.total_ms,
           "group_size": r.group_size, "upload_ms": r.upload_ms, "compute_ms": r.compute_ms,
           "followup_ms": r.followup_ms, "copy_ms": r.copy_ms,
           "vt_nnz": r.vt_nnz, "tt_nnz": r.tt_nnz}
    if r.vt_off:
        out["VT"] = (_arr(r.vt_off, V + 1, np.int64), _arr(r.vt_val, r.vt_nnz, np.int32))
    if r.tt_off:
        out["TT"] = (_arr(r.tt_off, T + 1, np.int64), _arr(r.tt_val, r.tt_nnz, np.int32))
    return out


class Server:
    def __init__(self, slots=2, hbm_budget_bytes=32 << 30, fuse_max_cells=1 << 20,
                 fuse_max_meshes=64):
        self._lib = _lib()
        self._h = C.c_void_p()
        self._keep = {}
        cfg = Config(slots, hbm_budget_bytes, fuse_max_cells, fuse_max_meshes)
        _native.check(self._lib.tg_server_create(C.byref(self._h), C.byref(cfg)))

    def submit(self, cells, n_points, scalars, followups=0) -> int:
        """followups: VT | TT to build with the first result (| COPY for host
        copies in the result)."""
        cells = np.ascontiguousarray(cells, np.int32).reshape(-1, 4)
        sc = np.ascontiguousarray(scalars, np.float32)
        rid = C.c_uint64()
        _native.check(self._lib.tg_server_submit(self._h, cells.ctypes.data, len(cells),
                                                 int(n_points), sc.ctypes.data, followups,
                                                 C.byref(rid)))
        self._keep[rid.value] = (cells, sc)  # inputs live until the first result
        return rid.value

    def wait(self, rid: int) -> dict:
        r = Result()
        _native.check(self._lib.tg_server_wait(self._h, rid, C.byref(r)))
        self._keep.pop(rid, None)
        return result_dict(r)

    def followup(self, rid: int, flags: int, copy: bool = True) -> dict:
        """VT/TT of the request, built on the device once and kept resident;
        copy=True also copies them to host memory.  copy=False returns the
        entry counts and the phase times only."""
        r = Result()
        _native.check(self._lib.tg_server_followup(self._h, rid, flags | (COPY if copy else 0),
                                                   C.byref(r)))
        if not copy:
            return {"vt_nnz": r.vt_nnz, "tt_nnz": r.tt_nnz, "followup_ms": r.followup_ms,
                    "copy_ms": r.copy_ms}
        return result_dict(r)

    def release(self, rid: int) -> None:
        _native.check(self._lib.tg_server_release(self._h, rid))

    def stats(self) -> dict:
        a, b, c = C.c_int64(), C.c_int64(), C.c_int64()
        _native.check(self._lib.tg_server_stats(self._h, C.byref(a), C.byref(b), C.byref(c)))
        return {"requests": a.value, "groups": b.value, "resident_bytes": c.value}

    def close(self) -> None:
        if self._h:
            self._lib.tg_server_destroy(self._h)
            self._h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
