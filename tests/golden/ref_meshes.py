"""The reference's own test meshes, regenerated without meshio.

artificial_dataset.sh:6-28 builds three meshes with make_mesh.py:
`--n-points 5 --output mwe`, `--n-points 100 --output tiny` (the reference's
default input file, include/datatypes.h.in fileName) and `--n-points 100000
--output big`, all with the default seed 1234 (make_mesh.py:15).

The fan rule (make_mesh.py:41-66), restated:
  - cell 0 is (0, 1, 2, 3); its four faces are the open faces;
  - the unused points n-1, n-2, ..., 4 are taken in that order (the list of
    set(range(n)) minus {0,1,2,3}, popped from its end);
  - each one takes an open face (Python's set.pop on the set of open faces,
    so the choice is CPython's), forms the cell sorted ascending, and the
    three new faces (each sorted) join the open set (a set union).
The scalar field is numpy's legacy generator seeded 1234 drawing rand(n)
(make_mesh.py:30 and :73) -- float64 in the .vtu file (the sed at :93 turns
the declared type into Float64).  Point positions are not needed by the
path and are not generated.

`python tests/golden/ref_meshes.py` rewrites ref_meshes.json (the full cells
and scalars of mwe and tiny; for big its sizes, SHA-256 digests of the int32
cells and float64 scalars, and a few sample cells) and ref_big_cells.npz
(big's cells: each set union copies the open set, so the rule is quadratic
and regenerating big takes minutes; tests load this file and regenerate
only a prefix of it).
"""
from __future__ import annotations

import hashlib
import json
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
FIXTURE = HERE / "ref_meshes.json"
BIG_CELLS = HERE / "ref_big_cells.npz"
MESHES = {"mwe": 5, "tiny": 100, "big": 100_000}


def fan_cells(n_points: int, max_cells: int | None = None) -> np.ndarray:
    """The rule's cells; max_cells stops after that many (a prefix)."""
    if n_points < 4:
        raise ValueError("a fan mesh needs at least 4 points")
    cells = [(0, 1, 2, 3)]
    todo = list(set(range(n_points)).difference({0, 1, 2, 3}))
    open_faces = {(0, 1, 2), (1, 2, 3), (0, 2, 3), (0, 1, 3)}
    steps = len(todo) if max_cells is None else min(len(todo), max_cells - 1)
    for _ in range(steps):
        p = todo.pop()
        a, b, c = open_faces.pop()
        cells.append(tuple(sorted((a, b, c, p))))
        fresh = {tuple(sorted((a, b, p))), tuple(sorted((b, c, p))), tuple(sorted((a, c, p)))}
        open_faces = open_faces.union(fresh)
    return np.asarray(cells, dtype=np.int32)


def fan_scalars(n_points: int, seed: int = 1234) -> np.ndarray:
    return np.random.RandomState(seed).rand(n_points).astype(np.float64)


def mesh(name: str):
    """(cells int32 (T, 4), n_points, scalars float64), regenerated."""
    n = MESHES[name]
    return fan_cells(n), n, fan_scalars(n)


def load(name: str):
    """(cells int32 (T, 4), n_points, scalars float64) from the committed
    files: mwe and tiny from the JSON fixture, big's cells from its .npz."""
    n = MESHES[name]
    if name == "big":
        with np.load(BIG_CELLS) as z:
            cells = np.ascontiguousarray(z["cells"], dtype=np.int32)
    else:
        cells = np.asarray(json.loads(FIXTURE.read_text())[name]["cells"], dtype=np.int32)
    return cells, n, fan_scalars(n)


def digest(a: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def fixture(big_cells: np.ndarray) -> dict:
    out = {}
    for name, n in MESHES.items():
        cells = big_cells if name == "big" else fan_cells(n)
        sc = fan_scalars(n)
        e = {"n_points": n, "n_cells": int(cells.shape[0]),
             "cells_sha256": digest(cells.astype("<i4")),
             "scalars_sha256": digest(sc.astype("<f8"))}
        if n <= 100:
            e["cells"] = cells.tolist()
            e["scalars"] = [float(x).hex() for x in sc]
        else:
            e["sample_cells"] = {str(i): cells[i].tolist() for i in (0, 1, 2, 1000, cells.shape[0] - 1)}
        out[name] = e
    return out


if __name__ == "__main__":
    big = fan_cells(MESHES["big"])
    FIXTURE.write_text(json.dumps(fixture(big), indent=1) + "\n")
    np.savez_compressed(BIG_CELLS, cells=big)
    print(f"wrote {FIXTURE} and {BIG_CELLS}")
