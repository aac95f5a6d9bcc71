"""PHAT ascii I/O and persistence diagrams.

The reference moves boundary matrices between MATLAB and PHAT/DIPHA as
ascii files (numerics/utils/dat2cmo.m:2-48, numerics/utils/cmo2dat.m:2-38,
numerics/utils/phat.m:46-79):

* a boundary file has one line per column, ``dim r_1 r_2 ...`` with 0-based
  row indices; lines starting with ``#`` and empty lines are skipped;
* a pairs file (what ``phat --ascii`` writes) holds the number of pairs on its
  first line, then one ``birth death`` line (0-based) per pair.

``read_boundary`` returns the 1-based COO the C ABI takes, with the vertex
columns left empty (dat2cmo.m marks them with row -1, and every caller drops
those entries, numerics/benchmark_datasets.m:17-18). ``write_boundary`` writes
the real dimension of each column, 0 for an empty one (cmo2dat.m writes -1
there; PHAT reads the dimension as given, and 0 is the dimension of a vertex).
"""
from __future__ import annotations

import os

import numpy as np

__all__ = ["read_boundary", "write_boundary", "read_pairs", "write_pairs", "pairs_from_low",
           "diagram"]


def read_boundary(path: str | os.PathLike):
    """Parse a PHAT ascii boundary file. Returns (rows1, cols1, m, dims):
    int32 1-based COO in column-major order with the rows of each column
    ascending, the number of columns, and the dimension written on each
    line."""
    rows, cols, dims = [], [], []
    m = 0
    with open(path) as f:
        for line in f:
            tok = line.split()
            if not tok or tok[0].startswith("#"):
                continue
            d = int(tok[0])
            entries = sorted(int(x) for x in tok[1:])
            if d < 0 or (d == 0) != (len(entries) == 0):
                raise ValueError(f"column {m}: dimension {d} with {len(entries)} faces")
            if any(r < 0 or r >= m for r in entries):
                raise ValueError(f"column {m}: a face index is not an earlier column")
            if len(set(entries)) != len(entries):
                raise ValueError(f"column {m}: repeated face")
            rows.extend(r + 1 for r in entries)
            cols.extend([m + 1] * len(entries))
            dims.append(d)
            m += 1
    return (np.asarray(rows, np.int32), np.asarray(cols, np.int32), m,
            np.asarray(dims, np.int32))


def write_boundary(path: str | os.PathLike, rows1, cols1, m: int, dims=None) -> None:
    """Write a 1-based column-major COO as a PHAT ascii boundary file. dims
    defaults to (number of faces - 1), and 0 for an empty column."""
    rows1 = np.asarray(rows1, np.int64)
    cols1 = np.asarray(cols1, np.int64)
    if rows1.shape != cols1.shape:
        raise ValueError("rows1 and cols1 differ in length")
    if cols1.size and (np.any(np.diff(cols1) < 0) or cols1.min() < 1 or cols1.max() > m):
        raise ValueError("cols1 must be non-decreasing in 1..m")
    start = np.searchsorted(cols1, np.arange(1, m + 2))
    with open(path, "w") as f:
        for j in range(m):
            faces = rows1[start[j]:start[j + 1]] - 1
            d = int(dims[j]) if dims is not None else max(len(faces) - 1, 0)
            f.write(" ".join([str(d)] + [str(int(x)) for x in faces]) + "\n")


def pairs_from_low(low1) -> np.ndarray:
    """(birth, death) pairs, 0-based, from a 1-based low array (0 = none),
    in increasing death order."""
    low1 = np.asarray(low1)
    death = np.flatnonzero(low1 > 0)
    return np.stack([low1[death] - 1, death], axis=1).astype(np.int64)


def write_pairs(path: str | os.PathLike, low1) -> None:
    """Write the pairs of a 1-based low array as PHAT does: the count, then
    one ``birth death`` line per pair (0-based)."""
    p = pairs_from_low(low1)
    with open(path, "w") as f:
        f.write(f"{len(p)}\n")
        for b, d in p:
            f.write(f"{b} {d}\n")


def read_pairs(path: str | os.PathLike, m: int) -> np.ndarray:
    """A PHAT pairs file back to a 1-based low array of m columns
    (phat.m:66-79: low(death + 1) = birth + 1)."""
    low1 = np.zeros(m, np.int32)
    with open(path) as f:
        for line in f:
            tok = line.split()
            if len(tok) != 2:
                continue
            b, d = int(tok[0]), int(tok[1])
            if not (0 <= b < d < m):
                raise ValueError(f"pair ({b}, {d}) is not inside 0..{m - 1} with birth < death")
            low1[d] = b + 1
    return low1


def diagram(low1, ess, dims, values=None) -> dict:
    """Persistence diagram per dimension: {dim: array of (birth, death)},
    with filtration values when given (else column indices), and death =
    inf for the essential classes. The dimension of a pair is that of its
    birth cell."""
    low1 = np.asarray(low1)
    ess = np.asarray(ess)
    dims = np.asarray(dims)
    v = np.arange(len(low1), dtype=np.float64) if values is None else np.asarray(values, np.float64)
    out = {}
    p = pairs_from_low(low1)
    for d in np.unique(dims):
        sel = p[dims[p[:, 0]] == d]
        finite = np.stack([v[sel[:, 0]], v[sel[:, 1]]], axis=1) if len(sel) else np.zeros((0, 2))
        e = np.flatnonzero((ess == 1) & (dims == d))
        inf = np.stack([v[e], np.full(len(e), np.inf)], axis=1) if len(e) else np.zeros((0, 2))
        out[int(d)] = np.concatenate([finite, inf], axis=0)
    return out
