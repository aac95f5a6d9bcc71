"""CPU oracle for the pms reduction - TEST INFRASTRUCTURE ONLY.

Only tests/, ``__graft_entry__.smoke()`` and bench.py's ``cpu_baseline`` leg
import this package, always as the checker; the product library
(``openph_amd``) never does.

* ``liboracle.so`` (built from ``oracle.c`` by ``oracle/Makefile``) holds
  the C restatements of the reference's serial ``standard``, ``twist``,
  ``ph_row`` and a serial ``pms`` (file:line citations in ``oracle.c``).
* ``dense_reduce`` is the dense Z/2 brute force that restates
  src/test/std_test.m:4-33 in numpy.
* ``dim0_pairs_vr`` gives the dimension-0 persistence pairs of a VR
  filtration by union-find with the elder rule (independent of any matrix
  reduction).
* ``ref_*`` call the reference's own ``reduce_col_host`` /
  ``left_to_right_host`` (src/cuda/algorithms/kernels_host.cu:2-75) compiled
  from /root/reference into ``oracle/_ref/libref.so`` by
  ``oracle/Makefile.ref``; available only where that build exists.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None
_REF = None

ORC_OK = 0
ORC_ENOMEM = -1
ORC_EOVERFLOW = -2


def build(force: bool = False) -> str:
    """Compile liboracle.so (gcc, a few seconds)."""
    so = os.path.join(_HERE, "liboracle.so")
    src = os.path.join(_HERE, "oracle.c")
    if force or not os.path.exists(so) or os.path.getmtime(so) < os.path.getmtime(src):
        subprocess.check_call(["make", "-C", _HERE, "-s", "liboracle.so"])
    return so


def build_ref() -> str | None:
    """Compile oracle/_ref/libref.so from the reference sources when they are
    present (build container only). Returns its path or None."""
    if not os.path.isdir("/root/reference/src/cuda"):
        p = os.path.join(_HERE, "_ref", "libref.so")
        return p if os.path.exists(p) else None
    subprocess.check_call(["make", "-C", _HERE, "-s", "-f", "Makefile.ref"])
    return os.path.join(_HERE, "_ref", "libref.so")


def _lib():
    global _LIB
    if _LIB is None:
        so = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(so):
            build()
        lib = ctypes.CDLL(so)
        i32p = np.ctypeslib.ndpointer(np.int32, flags="C")
        i64p = np.ctypeslib.ndpointer(np.int64, flags="C")
        for name in ("orc_standard", "orc_twist"):
            f = getattr(lib, name)
            f.restype = ctypes.c_int
            f.argtypes = [ctypes.c_int, i64p, i32p, ctypes.c_int, i32p, i32p, i64p]
        lib.orc_pms.restype = ctypes.c_int
        lib.orc_pms.argtypes = [ctypes.c_int, i64p, i32p, i32p, i32p, i32p, i64p,
                                ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                ctypes.c_void_p, ctypes.c_void_p]
        lib.orc_twist_dyn.restype = ctypes.c_int
        lib.orc_twist_dyn.argtypes = [ctypes.c_int, i64p, i32p, i32p, i32p, i64p]
        lib.orc_ph_row.restype = ctypes.c_int
        lib.orc_ph_row.argtypes = [ctypes.c_int, i64p, i32p, i32p, i32p, i64p, ctypes.c_int]
        _LIB = lib
    return _LIB


def _csc(b):
    colptr = np.ascontiguousarray(b.colptr, dtype=np.int64)
    rows = np.ascontiguousarray(b.rows, dtype=np.int32)
    return colptr, rows


class OracleResult:
    def __init__(self, low, ess, stats, num_iters=None, iter_stats=None):
        self.low = low
        self.ess = ess
        self.additions = int(stats[0])
        self.merge_elems = int(stats[1])
        self.max_len = int(stats[2])
        self.iters = int(stats[3])
        self.num_iters = num_iters
        self.iter_stats = iter_stats


def _check(rc, name):
    if rc != ORC_OK:
        raise RuntimeError(f"oracle {name} failed with code {rc}")


def standard(b, col_width: int = 1 << 20) -> OracleResult:
    colptr, rows = _csc(b)
    m = b.m
    low = np.empty(m, np.int32)
    ess = np.empty(m, np.int32)
    st = np.zeros(4, np.int64)
    cw = _safe_cw(b, col_width)
    _check(_lib().orc_standard(m, colptr, rows, cw, low, ess, st), "standard")
    return OracleResult(low, ess, st)


def twist(b, col_width: int = 1 << 20) -> OracleResult:
    colptr, rows = _csc(b)
    m = b.m
    low = np.empty(m, np.int32)
    ess = np.empty(m, np.int32)
    st = np.zeros(4, np.int64)
    cw = _safe_cw(b, col_width)
    _check(_lib().orc_twist(m, colptr, rows, cw, low, ess, st), "twist")
    return OracleResult(low, ess, st)


def twist_dyn(b) -> OracleResult:
    """twist on growable columns: orc_twist's order, clearing and merge with
    no slot limit (for matrices whose fixed-slot layout is too large)."""
    colptr, rows = _csc(b)
    m = b.m
    low = np.empty(m, np.int32)
    ess = np.empty(m, np.int32)
    st = np.zeros(4, np.int64)
    _check(_lib().orc_twist_dyn(m, colptr, rows, low, ess, st), "twist_dyn")
    return OracleResult(low, ess, st)


def _safe_cw(b, col_width):
    """Slots per column are col_width * max_nnz; clamp so m*p stays in a few
    hundred MB for the fixed-slot restatements."""
    mx = int(np.diff(b.colptr).max()) if b.m else 1
    mx = max(mx, 1)
    budget = (256 << 20) // 4  # ints
    cap = max(1, budget // max(1, b.m) // mx)
    return int(min(col_width, cap))


def pms(b, max_iters: int = 0, iter_stats: bool = False, snapshot: bool = False,
        low_true=None) -> OracleResult:
    """pms; with low_true (0-based, -1 = none) the result also carries `trk`,
    an (num_iters, 4) int64 array of the tracked integers per iteration:
    sum |low - low_true|, max |low - low_true|, #unreduced, #ess."""
    colptr, rows = _csc(b)
    m = b.m
    low = np.empty(m, np.int32)
    ess = np.empty(m, np.int32)
    st = np.zeros(4, np.int64)
    nit = np.zeros(1, np.int32)
    cap = 100000
    its = np.zeros(5 * cap, np.int64) if iter_stats else None
    ptr = its.ctypes.data if its is not None else None
    lt = None
    trk = None
    if low_true is not None:
        lt = np.ascontiguousarray(low_true, dtype=np.int32)
        if lt.shape[0] != m:
            raise ValueError("low_true must have m entries")
        trk = np.zeros(4 * (m + 1), np.int64)
    _check(_lib().orc_pms(m, colptr, rows, low, ess, nit, st, ptr, cap, max_iters,
                          int(snapshot), lt.ctypes.data if lt is not None else None,
                          trk.ctypes.data if trk is not None else None), "pms")
    per = None
    if its is not None:
        per = its.reshape(-1, 5)[: int(st[3])].copy()
    res = OracleResult(low, ess, st, int(nit[0]), per)
    res.trk = trk.reshape(-1, 4)[: int(nit[0])].copy() if trk is not None else None
    return res


def ph_row(b, parallel: bool = False) -> OracleResult:
    """ph_row, or ph_row_parallel with ``parallel`` (every later column with
    the pivot's low is a neighbour; ess from ess_hat after every row)."""
    colptr, rows = _csc(b)
    m = b.m
    low = np.empty(m, np.int32)
    ess = np.empty(m, np.int32)
    st = np.zeros(4, np.int64)
    _check(_lib().orc_ph_row(m, colptr, rows, low, ess, st, 1 if parallel else 0), "ph_row")
    return OracleResult(low, ess, st)


# ---------------------------------------------------------------------------
# Independent checks
# ---------------------------------------------------------------------------

def dense_reduce(b):
    """Dense Z/2 standard reduction (restates src/test/std_test.m:4-33):
    returns (low, ess) 0-based. Only for m up to a few thousand."""
    m = b.m
    M = np.zeros((m, m), dtype=np.uint8)
    for j in range(m):
        M[b.rows[b.colptr[j]:b.colptr[j + 1]], j] = 1
    low = np.full(m, -1, np.int64)
    owner = {}
    for j in range(m):
        col = M[:, j]
        while True:
            nz = np.flatnonzero(col)
            if nz.size == 0:
                low[j] = -1
                break
            lj = int(nz[-1])
            if lj in owner:
                col ^= M[:, owner[lj]]
            else:
                low[j] = lj
                owner[lj] = j
                break
        M[:, j] = col
    ess = np.ones(m, np.int32)
    for j in range(m):
        if low[j] >= 0:
            ess[j] = 0
            ess[low[j]] = 0
    return low.astype(np.int32), ess


def dim0_pairs_vr(b):
    """Dimension-0 pairs of a filtration whose dimension-1 cells are edges:
    union-find over edges in filtration order with the elder rule (the
    component whose oldest vertex is younger dies). Returns {edge: vertex}."""
    parent = {}

    def find(x):
        while parent[x] != x:
            parent[x] = parent[parent[x]]
            x = parent[x]
        return x

    pairs = {}
    for j in range(b.m):
        if b.dim[j] == 0:
            parent[j] = j
        elif b.dim[j] == 1:
            u, v = (int(x) for x in b.rows[b.colptr[j]:b.colptr[j + 1]])
            ru, rv = find(u), find(v)
            if ru != rv:
                young, old = (ru, rv) if ru > rv else (rv, ru)
                pairs[j] = young
                parent[young] = old
    return pairs


def pairs_by_dim(low, ess, dim):
    """Persistence pairs grouped by the dimension of the birth cell, plus
    essential cells per dimension."""
    out = {}
    for j in np.flatnonzero(low >= 0):
        d = int(dim[low[j]])
        out.setdefault(d, []).append((int(low[j]), int(j)))
    esses = {}
    for j in np.flatnonzero(ess == 1):
        esses.setdefault(int(dim[j]), []).append(int(j))
    return out, esses


# ---------------------------------------------------------------------------
# The reference's own merge, compiled from /root/reference (build container)
# ---------------------------------------------------------------------------

def _ref():
    global _REF
    if _REF is None:
        so = os.path.join(_HERE, "_ref", "libref.so")
        if not os.path.exists(so):
            return None
        lib = ctypes.CDLL(so)
        i32p = np.ctypeslib.ndpointer(np.int32, flags="C")
        i64p = np.ctypeslib.ndpointer(np.int64, flags="C")
        for name in ("ref_standard", "ref_twist"):
            f = getattr(lib, name)
            f.restype = ctypes.c_int
            f.argtypes = [ctypes.c_int, i64p, i32p, ctypes.c_int, i32p, i32p]
        _REF = lib
    return _REF


def ref_available() -> bool:
    return _ref() is not None


def ref_run(name: str, b, col_width: int):
    """Run the reference's reduce_col_host in standard or twist order. The
    reference exits the process on slot overflow, so callers run this in a
    subprocess or choose col_width generously."""
    lib = _ref()
    colptr, rows = _csc(b)
    low = np.empty(b.m, np.int32)
    ess = np.empty(b.m, np.int32)
    rc = getattr(lib, "ref_" + name)(b.m, colptr, rows, col_width, low, ess)
    _check(rc, "ref_" + name)
    return low, ess
