"""Critical-point labels against a TTK / ParaView export.

The reference checks its classes against TTK's ScalarFieldCriticalPoints
through two scripts: convert_vtk_to_txt.py:8-59 writes one "<index> <class>"
line per critical point (with --full-vtk the index is the point's global
index; the class is TTK's CriticalType integer or, with --class-as-string,
its name), and pyclassify.py:63-126 reads that file, treats every point
without a line as regular, maps TTK's six types onto the four numeric
classes and reports the points it disagrees on.

Here the same comparison works on the six-way labels (`Mesh.labels6`, TTK's
numbering: 0 minimum, 1 1-saddle, 2 2-saddle, 3 maximum, 4 multi-saddle,
5 regular) and on the four numeric classes (`Mesh.classes()[:, 2]`:
1 minimum, 2 maximum, 3 regular, 4 saddle).  TTK itself is not part of this
build, so agreement with TTK's own output is only as good as the export
given.
"""
from __future__ import annotations

from pathlib import Path

import numpy as np

TTK_NAMES = ("Minima", "1-Saddle", "2-Saddle", "Maxima", "Multi-Saddle", "Regular")
LABEL6_OF_NAME = {n: i for i, n in enumerate(TTK_NAMES)}
# TTK type -> numeric class (1 minimum, 2 maximum, 3 regular, 4 saddle); the
# same grouping as pyclassify.py:65 (saddles of every kind are one class)
CLASS4_OF_TTK = np.array([1, 4, 4, 2, 4, 3], np.uint8)


def read_export(path, n_points: int) -> np.ndarray:
    """TTK types (uint8, n_points) from an "<index> <class>" export; points
    without a line are regular (5).  A line that is not exactly two fields, a
    class that is neither 0..5 nor a TTK name, or an index outside
    [0, n_points) is an error."""
    types = np.full(n_points, 5, np.uint8)
    for no, line in enumerate(Path(path).read_text().splitlines(), 1):
        if not line.strip():
            continue
        parts = line.split()
        if len(parts) != 2:
            raise ValueError(f"{path}:{no}: expected '<index> <class>', got {line!r}")
        idx = int(parts[0])
        cls = parts[1]
        k = LABEL6_OF_NAME.get(cls)
        if k is None:
            k = int(cls)
            if not 0 <= k <= 5:
                raise ValueError(f"{path}:{no}: class {k} is not a TTK critical type")
        if not 0 <= idx < n_points:
            raise ValueError(f"{path}:{no}: index {idx} outside [0, {n_points})")
        types[idx] = k
    return types


def write_export(path, labels6, as_string: bool = False) -> int:
    """The export format for labels6: one line per non-regular point
    (regular points are not logged).  Returns the number of lines."""
    lab = np.asarray(labels6, np.uint8)
    idx = np.flatnonzero(lab != 5)
    with open(path, "w") as f:
        for i in idx:
            f.write(f"{i} {TTK_NAMES[lab[i]] if as_string else int(lab[i])}\n")
    return int(idx.size)


def compare(labels6, ttk_types, classes=None, max_listed: int = 10) -> dict:
    """Agreement of our labels with TTK's types.

    Six-way: labels6 against ttk_types, with the 6x6 confusion matrix
    (rows ours, columns TTK) and the first disagreeing points.  Four-way
    (with `classes`, the (V, 3) array or its class column): the numeric
    class against CLASS4_OF_TTK of TTK's type, as pyclassify.py compares."""
    lab = np.asarray(labels6, np.uint8)
    ttk = np.asarray(ttk_types, np.uint8)
    if lab.shape != ttk.shape:
        raise ValueError(f"{lab.shape[0]} labels against {ttk.shape[0]} TTK points")
    conf = np.zeros((6, 6), np.int64)
    np.add.at(conf, (lab, ttk), 1)
    bad6 = np.flatnonzero(lab != ttk)
    out = {"points": int(lab.size), "agree6": int(lab.size - bad6.size),
           "mismatch6": int(bad6.size), "first_mismatch6": bad6[:max_listed].tolist(),
           "confusion6": conf.tolist(),
           "counts_ours": {TTK_NAMES[k]: int((lab == k).sum()) for k in range(6)},
           "counts_ttk": {TTK_NAMES[k]: int((ttk == k).sum()) for k in range(6)}}
    if classes is not None:
        c = np.asarray(classes)
        c = c[:, 2] if c.ndim == 2 else c
        bad4 = np.flatnonzero(c.astype(np.uint8) != CLASS4_OF_TTK[ttk])
        out.update({"mismatch4": int(bad4.size), "first_mismatch4": bad4[:max_listed].tolist()})
    return out
