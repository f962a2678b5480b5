"""ctypes bindings of the native libraries (libtetra_hip.so, libtetra_host.so).

The libraries are built in-tree by tetragpu_amd/build.py.  Loading fails
loudly when they are missing: there is no Python or CPU fallback for any
device computation.
"""
from __future__ import annotations

import ctypes as C
import re
from pathlib import Path

LIB_DIR = Path(__file__).resolve().parent / "lib"
INCLUDE_DIR = Path(__file__).resolve().parent.parent / "include"


class NativeError(RuntimeError):
    """A native call returned a TG_ERR_* status."""

    def __init__(self, status: int, message: str):
        super().__init__(f"[{status}] {message}")
        self.status = status


_hip = None
_host = None

P = C.c_void_p
I32, I64, U32 = C.c_int32, C.c_int64, C.c_uint32

_HIP_SIGS = {
    "tg_init": (C.c_int, [C.c_int]),
    "tg_last_error": (C.c_char_p, []),
    "tg_device_count": (C.c_int, [C.POINTER(C.c_int)]),
    "tg_mesh_create": (C.c_int, [C.POINTER(P), P, I64, I64, P, C.c_int, C.c_int, P]),
    "tg_mesh_destroy": (C.c_int, [P]),
    "tg_mesh_set_scalars": (C.c_int, [P, P, C.c_int, C.c_int]),
    "tg_mesh_stream": (C.c_int, [P, C.POINTER(P)]),
    "tg_mesh_counts": (C.c_int, [P, C.POINTER(I64), C.POINTER(I64), C.POINTER(I64),
                                 C.POINTER(I64)]),
    "tg_mesh_set_config": (C.c_int, [P, I32, I32, I32]),
    "tg_build": (C.c_int, [P, U32]),
    "tg_critpoints": (C.c_int, [P]),
    "tg_vv_critpoints": (C.c_int, [P]),
    "tg_sync": (C.c_int, [P]),
    "tg_relation_shape": (C.c_int, [P, I32, C.POINTER(I64), C.POINTER(I64), C.POINTER(I32),
                                    C.POINTER(I32)]),
    "tg_relation_copy": (C.c_int, [P, I32, P, P]),
    "tg_relation_device": (C.c_int, [P, I32, C.POINTER(P), C.POINTER(I32), C.POINTER(P)]),
    "tg_classes_copy": (C.c_int, [P, P]),
    "tg_classes_device": (C.c_int, [P, C.POINTER(P)]),
    "tg_labels6_copy": (C.c_int, [P, P]),
    "tg_approx_max_vv": (C.c_int, [P, C.POINTER(I32)]),
    "tg_device_free": (C.c_int, [P]),
    "tg_relation_export": (C.c_int, [P, I32, I32, C.POINTER(P), I32, C.POINTER(P)]),
    "tg_device_upload": (C.c_int, [P, C.c_size_t, C.POINTER(P)]),
    "tg_mesh_set_bucket": (C.c_int, [P, I32, I64]),
    "tg_mesh_set_routing": (C.c_int, [P, I32]),
    "tg_mesh_routing": (C.c_int, [P, C.POINTER(I32)]),
    "tg_set_tuning": (C.c_int, [P, I32]),
    "tg_get_tuning": (C.c_int, [P, I32, C.POINTER(I32)]),
    "tg_batch_create": (C.c_int, [C.POINTER(P), I32, P, P, P, P, C.c_int, I64, P]),
    "tg_batch_destroy": (C.c_int, [P]),
    "tg_batch_counts": (C.c_int, [P, C.POINTER(I32), C.POINTER(I64), C.POINTER(I64)]),
    "tg_batch_use_graph": (C.c_int, [P, C.c_int]),
    "tg_batch_set_config": (C.c_int, [P, I32, I32, I32, I32, I64]),
    "tg_batch_vv_critpoints": (C.c_int, [P]),
    "tg_batch_sync": (C.c_int, [P]),
    "tg_batch_classes_copy": (C.c_int, [P, P]),
    "tg_batch_vv_copy": (C.c_int, [P, I32, P, P, C.POINTER(I64)]),
    "tg_batch_build": (C.c_int, [P, U32]),
    "tg_stream_create": (C.c_int, [C.POINTER(P)]),
    "tg_pool_trim": (C.c_int, []),
    "tg_host_alloc": (C.c_int, [C.c_size_t, C.POINTER(P)]),
    "tg_host_free": (C.c_int, [P]),
    "tg_pool_cached": (C.c_int, [C.POINTER(I64)]),
    "tg_stream_destroy": (C.c_int, [P]),
    "tg_batch_relation_copy": (C.c_int, [P, I32, I32, P, P, C.POINTER(I64)]),
    "tg_profile_enable": (C.c_int, [P, C.c_int]),
    "tg_profile_read": (C.c_int, [P, I32, C.POINTER(I32), P, P]),
    "tg_vv_critpoints_noalloc": (C.c_int, [P, I64, I64, P, C.c_int, P, P, I64, P, P, P]),
    "tg_noalloc_limits": (C.c_int, [I64, I64, C.POINTER(I64), C.POINTER(I64)]),
    "tg_vv_footprint": (C.c_int, [I64, I64, C.POINTER(I64)]),
    "tg_device_budget": (C.c_int, [C.POINTER(I64)]),
    "tg_critpoints_split": (C.c_int, [P, I64, I64, P, C.c_int, P, I32, P, I64, P, P]),
}

_HOST_SIGS = {
    "tg_gen_kuhn": (C.c_int, [I32, C.c_uint64, U32, C.c_double, I32, I32, I32, P]),
    "tg_gen_scalars": (C.c_int, [P, I32, C.c_uint64]),
    "tg_gen_free": (None, [P]),
    "tg_autotune": (C.c_int, [P, I32, I32, P, I32, C.POINTER(I32)]),
    "tg_tuning_load": (C.c_int, [C.c_char_p]),
    "tg_tuning_save": (C.c_int, [C.c_char_p]),
    "tetra_partitioned_classes": (C.c_int, [P, C.c_int64, C.c_int64, P, P, P, P]),
    "tetra_partition_error": (C.c_char_p, []),
}


def _load(name: str) -> C.CDLL:
    path = LIB_DIR / name
    if not path.exists():
        raise ImportError(f"{path} is missing; build it with `python -m tetragpu_amd.build`")
    return C.CDLL(str(path), mode=C.RTLD_GLOBAL)


def _bind(lib: C.CDLL, sigs: dict) -> None:
    for fn, (res, args) in sigs.items():
        f = getattr(lib, fn)
        f.restype = res
        f.argtypes = args


def hip() -> C.CDLL:
    global _hip
    if _hip is None:
        lib = _load("libtetra_hip.so")
        _bind(lib, _HIP_SIGS)
        _hip = lib
    return _hip


def host() -> C.CDLL:
    global _host
    if _host is None:
        hip()  # libtetra_host.so links against it
        lib = _load("libtetra_host.so")
        _bind(lib, _HOST_SIGS)
        _host = lib
    return _host


def check(status: int) -> None:
    if status != 0:
        raise NativeError(status, hip().tg_last_error().decode(errors="replace"))


def declared_symbols() -> dict[str, list[str]]:
    """Function names declared in include/*.h, per header."""
    out = {}
    for h in sorted(INCLUDE_DIR.glob("*.h")):
        text = re.sub(r"/\*.*?\*/", "", h.read_text(), flags=re.S)
        names = re.findall(r"\b(tg_[a-z0-9_]+)\s*\(", text)
        out[h.name] = sorted(set(n for n in names if not n.isupper()))
    return out
