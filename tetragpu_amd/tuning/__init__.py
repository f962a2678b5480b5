"""Autotuning of the kernel form and launch shape per mesh size class.

`autotune(sides)` times every candidate on generated Kuhn grids of the given
sides (on the GPU), installs the fastest per size class and returns the
table; `load()` installs a saved table (default: the one measured on
MI355X, `mi355x.json` next to this file)."""
from __future__ import annotations

import ctypes as C
import json
from pathlib import Path

import numpy as np

from .. import _native
from .._native import check

DEFAULT = Path(__file__).resolve().parent / "mi355x.json"
FIELDS = ["max_cells", "variant", "lane_cap", "lane_block", "bv_log", "tets_per_chunk"]


class TuneEntry(C.Structure):
    _fields_ = [("max_cells", C.c_int64), ("variant", C.c_int32), ("lane_cap", C.c_int32),
                ("lane_block", C.c_int32), ("bv_log", C.c_int32), ("tets_per_chunk", C.c_int64)]


def autotune(sides=(6, 10, 16, 24, 40, 64, 96, 128), reps: int = 5) -> list[dict]:
    lib = _native.host()
    s = np.ascontiguousarray(sides, np.int32)
    out = (TuneEntry * 64)()
    n = C.c_int32()
    check(lib.tg_autotune(s.ctypes.data, len(s), reps, out, 64, C.byref(n)))
    return [{f: getattr(out[i], f) for f in FIELDS} for i in range(min(n.value, 64))]


def install(table: list[dict]) -> None:
    arr = (TuneEntry * len(table))(*[TuneEntry(*[int(e[f]) for f in FIELDS]) for e in table])
    check(_native.hip().tg_set_tuning(arr, len(table)))


def current() -> list[dict]:
    out = (TuneEntry * 64)()
    n = C.c_int32()
    check(_native.hip().tg_get_tuning(out, 64, C.byref(n)))
    return [{f: getattr(out[i], f) for f in FIELDS} for i in range(min(n.value, 64))]


def load(path: str | Path = DEFAULT) -> list[dict] | None:
    p = Path(path)
    if not p.exists():
        return None
    table = json.loads(p.read_text())
    install(table)
    return table


def save(table: list[dict], path: str | Path) -> None:
    Path(path).write_text(json.dumps(table, indent=1) + "\n")
