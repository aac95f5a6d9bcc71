"""openph_amd - MI355X-native pms boundary-matrix reduction.

Python mirror of the reference's MATLAB interface over the C ABI in
``include/openph.h`` (``libopenph.so``, built from ``openph_amd/csrc``):

* ``ph(algstr, r, c, m, col_width, low_true)`` returns the same eight
  outputs as the mex gateway ``ph`` (src/matlab/ph/ph.cu:11-92):
  ``low, ess, err_linf, err_lone, err_redu, err_ess, time_track, num_iters``.
* ``openph(r, c, m, algo, col_width, low_true)`` returns ``(OUT, t)`` like
  src/matlab/ph/openph.m:3-21, OUT being a dict with the same fields.
* ``pms_device`` / ``pms_batch`` / ``pms_batch_device`` expose the device
  resident and batched entry points.

The library is required: importing works without it, but every call raises
``RuntimeError`` when ``libopenph.so`` is missing or cannot be loaded.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
import time

import numpy as np

__all__ = [
    "OpenPHError", "Options", "Stats", "lib", "build", "ph", "openph", "reduce",
    "pms_device", "pms_batch", "pms_batch_device", "autotune_batch_device", "TuneRow",
    "vr2_device", "cubical_device", "ALGORITHMS",
]

_HERE = os.path.dirname(os.path.abspath(__file__))
_SO = os.path.join(_HERE, "libopenph.so")
_LIB = None

ALGORITHMS = ("pms", "standard", "twist", "ph_row",
              "standard_parallel", "twist_parallel", "ph_row_parallel")


class OpenPHError(RuntimeError):
    def __init__(self, status: int, what: str = ""):
        self.status = status
        msg = lib().openph_strerror(status).decode() if _LIB is not None else str(status)
        super().__init__(f"{what}: {msg} (status {status})" if what else f"{msg} (status {status})")


class Options(ctypes.Structure):
    """openph_options (include/openph.h)."""
    _fields_ = [
        ("device", ctypes.c_int),
        ("max_iters", ctypes.c_int),
        ("pool_elems", ctypes.c_int64),
        ("pool_max_elems", ctypes.c_int64),
        ("lanes_per_column", ctypes.c_int),
        ("track", ctypes.c_int),
        ("slot_elems", ctypes.c_int),
        ("group_lds", ctypes.c_int),
        ("schedule", ctypes.c_int),
        ("chunk_iters", ctypes.c_int),
        ("time_phase2", ctypes.c_int),
    ]

    # values of ``schedule``
    AUTO, SYNC, QUEUED, GRAPH = 0, 1, 2, 3


class Stats(ctypes.Structure):
    """openph_stats (include/openph.h)."""
    _fields_ = [
        ("num_iters", ctypes.c_int),
        ("compactions", ctypes.c_int),
        ("additions", ctypes.c_int64),
        ("merge_elems", ctypes.c_int64),
        ("pool_peak", ctypes.c_int64),
        ("device_ms", ctypes.c_float),
        ("phase2_ms", ctypes.c_float),
        ("phase2_launches", ctypes.c_int),
        ("m", ctypes.c_int),
        ("overflows", ctypes.c_int),
        ("lanes", ctypes.c_int),
        ("spills", ctypes.c_int),
        ("group_lds", ctypes.c_int),
        ("handoffs", ctypes.c_int),
        ("schedule", ctypes.c_int),
        ("host_reads", ctypes.c_int),
    ]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class TuneRow(ctypes.Structure):
    """openph_tune_row (include/openph.h)."""
    _fields_ = [
        ("lanes_per_column", ctypes.c_int),
        ("group_lds", ctypes.c_int),
        ("slot_elems", ctypes.c_int),
        ("schedule", ctypes.c_int),
        ("wall_ms", ctypes.c_float),
        ("same", ctypes.c_int),
    ]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


def build(force: bool = True) -> str:
    """Compile libopenph.so for gfx950 with hipcc (in-tree)."""
    args = ["make", "-C", os.path.join(_HERE, "csrc"), "-j8"]
    if force:
        args.append("-B")
    subprocess.check_call(args)
    return _SO


def lib():
    """Load libopenph.so; raise if it is missing (there is no fallback)."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(_SO):
            raise RuntimeError(f"{_SO} is not built; run openph_amd.build() or "
                               "__graft_entry__.build() first")
        L = ctypes.CDLL(_SO)
        c_int_p = ctypes.POINTER(ctypes.c_int)
        c_float_p = ctypes.POINTER(ctypes.c_float)
        L.openph.restype = None
        L.openph.argtypes = [ctypes.c_char_p, c_int_p, c_int_p, ctypes.c_int, ctypes.c_int,
                             c_int_p, ctypes.c_int, c_int_p, c_int_p, c_float_p, c_float_p,
                             c_float_p, c_float_p, c_float_p, c_int_p, ctypes.c_int]
        L.openph_v2.restype = ctypes.c_int
        L.openph_v2.argtypes = [ctypes.c_char_p, c_int_p, c_int_p, ctypes.c_int, ctypes.c_int,
                                c_int_p, ctypes.POINTER(Options), c_int_p, c_int_p, c_float_p,
                                ctypes.POINTER(Stats)]
        L.openph_pms_device.restype = ctypes.c_int
        L.openph_pms_device.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int,
                                        ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
                                        ctypes.c_void_p, ctypes.POINTER(Options),
                                        ctypes.POINTER(Stats)]
        L.openph_pms_device_noalloc.restype = ctypes.c_int
        L.openph_pms_device_noalloc.argtypes = L.openph_pms_device.argtypes
        L.openph_pms_batch.restype = ctypes.c_int
        L.openph_pms_batch.argtypes = [ctypes.c_int, c_int_p, c_int_p, c_int_p, c_int_p,
                                       ctypes.POINTER(Options), c_int_p, c_int_p,
                                       ctypes.POINTER(Stats)]
        L.openph_pms_batch_device.restype = ctypes.c_int
        L.openph_pms_batch_device.argtypes = [ctypes.c_int, c_int_p, c_int_p, ctypes.c_void_p,
                                              ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                              ctypes.c_void_p, ctypes.POINTER(Options),
                                              ctypes.POINTER(Stats)]
        L.openph_autotune_batch_device.restype = ctypes.c_int
        L.openph_autotune_batch_device.argtypes = [
            ctypes.c_int, c_int_p, c_int_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
            ctypes.c_void_p, ctypes.c_void_p, ctypes.POINTER(Options), ctypes.c_int,
            ctypes.POINTER(Options), ctypes.POINTER(TuneRow), ctypes.c_int,
            ctypes.POINTER(ctypes.c_int)]
        i64p = ctypes.POINTER(ctypes.c_int64)
        L.openph_vr2_size.restype = ctypes.c_int
        L.openph_vr2_size.argtypes = [ctypes.c_int, i64p, i64p]
        L.openph_vr2_device.restype = ctypes.c_int
        L.openph_vr2_device.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int,
                                        ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                        ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int]
        L.openph_cubical_size.restype = ctypes.c_int
        L.openph_cubical_size.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, i64p, i64p]
        L.openph_cubical_device.restype = ctypes.c_int
        L.openph_cubical_device.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int,
                                            ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
                                            ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                            ctypes.c_int]
        L.openph_strerror.restype = ctypes.c_char_p
        L.openph_strerror.argtypes = [ctypes.c_int]
        L.openph_version.restype = ctypes.c_char_p
        L.openph_version.argtypes = []
        _LIB = L
    return _LIB


def _iptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))


def _fptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


def _i32(x):
    return np.ascontiguousarray(np.asarray(x).reshape(-1), dtype=np.int32)


def ph(algstr, r, c, m, col_width=1, low_true=None, device: int = 0, schedule: int = 0,
       options: Options | None = None):
    """The mex gateway's contract (src/matlab/ph/ph.cu:11-92): 1-based COO in,
    eight outputs back. Errors raise OpenPHError instead of printing.

    low_true (1-based, 0 = none) feeds the err_* trackers; without it they
    keep the gateway's initial -1. schedule selects how the host drives the
    pms iterations (Options.schedule; 0 = automatic, as the drop-in does).
    options, when given, supplies the other fields (pool, lane shape)."""
    rows = _i32(r)
    cols = _i32(c)
    m = int(m)
    nnz = rows.shape[0]
    if cols.shape[0] != nnz:
        raise OpenPHError(-1, "rows and cols differ in length")
    lt = _i32(low_true) if low_true is not None else None
    if lt is not None and lt.shape[0] < m:
        raise OpenPHError(-1, "low_true has fewer than m entries")
    # outputs initialised as the gateway does (ph.cu:72-78)
    low = np.full(m, -1, np.int32)
    ess = np.full(m, -1, np.int32)
    trk = np.full((5, m + 1), -1.0, np.float32)
    trk[4, :] = 0.0
    opt = Options()
    if options is not None:
        ctypes.memmove(ctypes.byref(opt), ctypes.byref(options), ctypes.sizeof(Options))
    opt.device = device
    opt.track = 1
    opt.schedule = int(schedule)
    st = Stats()
    rc = lib().openph_v2(algstr.encode(), _iptr(rows), _iptr(cols), m, nnz,
                         _iptr(lt) if lt is not None else None, ctypes.byref(opt), _iptr(low),
                         _iptr(ess), _fptr(trk), ctypes.byref(st))
    if rc != 0:
        raise OpenPHError(rc, f"ph('{algstr}')")
    return (low, ess, trk[0], trk[1], trk[2], trk[3], trk[4],
            np.array([st.num_iters], np.int32))


def openph(r, c, m, algo, col_width=1, low_true=None):
    """src/matlab/ph/openph.m:3-21: returns (OUT, t) with OUT a dict."""
    t0 = time.perf_counter()
    o = ph(algo, r, c, m, col_width, low_true)
    t = time.perf_counter() - t0
    out = {
        "low": o[0].astype(np.float64), "ess": o[1].astype(np.float64),
        "err_linf": o[2].astype(np.float64), "err_lone": o[3].astype(np.float64),
        "err_redu": o[4].astype(np.float64), "err_ess": o[5].astype(np.float64),
        "time_track": o[6].astype(np.float64), "num_iters": float(o[7][0]),
    }
    n = int(o[7][0])
    out["time_inner"] = float(out["time_track"][:n].sum())
    out["m"] = float(m)
    return out, t


def reduce(boundary, algstr: str = "pms", options: Options | None = None):
    """Reduce a ``complexes.Boundary``; returns (low0, ess, Stats) with low
    0-based (-1 = zero column)."""
    rows, cols = boundary.coo1()
    m = boundary.m
    low = np.full(m, -1, np.int32)
    ess = np.full(m, -1, np.int32)
    st = Stats()
    opt = options or Options()
    rc = lib().openph_v2(algstr.encode(), _iptr(rows), _iptr(cols), m, rows.shape[0], None,
                         ctypes.byref(opt), _iptr(low), _iptr(ess), None, ctypes.byref(st))
    if rc != 0:
        raise OpenPHError(rc, f"reduce('{algstr}')")
    return low - 1, ess, st


def pms_device(rows1, cols1, m, low1, ess, stream=None, options: Options | None = None):
    """Device-resident pms: rows1/cols1/low1/ess are int32 torch tensors (or
    raw device addresses) on the device of ``options.device``. Returns Stats."""
    def addr(x):
        return x if isinstance(x, int) else x.data_ptr()
    nnz = rows1.numel() if hasattr(rows1, "numel") else None
    if nnz is None:
        raise ValueError("pass torch tensors for rows1/cols1")
    st = Stats()
    opt = options or Options()
    s = stream.cuda_stream if stream is not None and hasattr(stream, "cuda_stream") else stream
    rc = lib().openph_pms_device(addr(rows1), addr(cols1), int(m), int(nnz), addr(low1),
                                 addr(ess), s, ctypes.byref(opt), ctypes.byref(st))
    if rc != 0:
        raise OpenPHError(rc, "pms_device")
    return st


def pms_batch(boundaries, options: Options | None = None):
    """Reduce many matrices in one batched engine pass. Returns a list of
    (low0, ess) per matrix and the Stats of the batch."""
    ms = np.array([b.m for b in boundaries], np.int32)
    nnzs = np.array([b.nnz for b in boundaries], np.int32)
    rr, cc = [], []
    for b in boundaries:
        r, c = b.coo1()
        rr.append(r)
        cc.append(c)
    rows = np.ascontiguousarray(np.concatenate(rr) if rr else np.zeros(0, np.int32))
    cols = np.ascontiguousarray(np.concatenate(cc) if cc else np.zeros(0, np.int32))
    M = int(ms.sum())
    low = np.full(M, -1, np.int32)
    ess = np.full(M, -1, np.int32)
    st = Stats()
    opt = options or Options()
    rc = lib().openph_pms_batch(len(boundaries), _iptr(ms), _iptr(nnzs), _iptr(rows),
                                _iptr(cols), ctypes.byref(opt), _iptr(low), _iptr(ess),
                                ctypes.byref(st))
    if rc != 0:
        raise OpenPHError(rc, "pms_batch")
    out = []
    o = 0
    for b in boundaries:
        out.append((low[o:o + b.m] - 1, ess[o:o + b.m]))
        o += b.m
    return out, st


def pms_batch_device(ms, nnzs, rows1, cols1, low1, ess, stream=None,
                     options: Options | None = None):
    """Batched pms on device tensors (concatenated local 1-based COO)."""
    ms = _i32(ms)
    nnzs = _i32(nnzs)
    st = Stats()
    opt = options or Options()
    s = stream.cuda_stream if stream is not None and hasattr(stream, "cuda_stream") else stream
    rc = lib().openph_pms_batch_device(len(ms), _iptr(ms), _iptr(nnzs), rows1.data_ptr(),
                                       cols1.data_ptr(), low1.data_ptr(), ess.data_ptr(), s,
                                       ctypes.byref(opt), ctypes.byref(st))
    if rc != 0:
        raise OpenPHError(rc, "pms_batch_device")
    return st


def autotune_batch_device(ms, nnzs, rows1, cols1, low1, ess, stream=None,
                          base: Options | None = None, reps: int = 3):
    """Autotune the batched path on device tensors (arguments as
    ``pms_batch_device``). Returns (best Options, list of measured rows as
    dicts); every candidate's output is checked against the first's."""
    ms = _i32(ms)
    nnzs = _i32(nnzs)
    best = Options()
    rows = (TuneRow * 64)()
    n = ctypes.c_int(0)
    s = stream.cuda_stream if stream is not None and hasattr(stream, "cuda_stream") else stream
    rc = lib().openph_autotune_batch_device(
        len(ms), _iptr(ms), _iptr(nnzs), rows1.data_ptr(), cols1.data_ptr(), low1.data_ptr(),
        ess.data_ptr(), s, ctypes.byref(base or Options()), int(reps), ctypes.byref(best), rows,
        64, ctypes.byref(n))
    if rc != 0:
        raise OpenPHError(rc, "autotune_batch_device")
    return best, [rows[i].as_dict() for i in range(n.value)]


def _stream_ptr(stream):
    return stream.cuda_stream if stream is not None and hasattr(stream, "cuda_stream") else stream


def vr2_device(points, stream=None, device: int = 0):
    """Full-radius VR 2-skeleton of an (n, d) float64 CUDA tensor of points
    (d <= 7), built on the device. Returns (rows1, cols1, values, dims) as
    device tensors: the 1-based COO in filtration order, each column's value
    and dimension (the same matrix as complexes.vr_2skeleton)."""
    import torch

    n, d = points.shape
    m = ctypes.c_int64()
    nnz = ctypes.c_int64()
    lib().openph_vr2_size(n, ctypes.byref(m), ctypes.byref(nnz))
    dev = points.device
    rows = torch.empty(max(nnz.value, 1), dtype=torch.int32, device=dev)
    cols = torch.empty_like(rows)
    vals = torch.empty(max(m.value, 1), dtype=torch.float64, device=dev)
    dims = torch.empty(max(m.value, 1), dtype=torch.int32, device=dev)
    pts = points.contiguous()
    rc = lib().openph_vr2_device(pts.data_ptr(), n, d, rows.data_ptr(), cols.data_ptr(),
                                 vals.data_ptr(), dims.data_ptr(), _stream_ptr(stream), device)
    if rc != 0:
        raise OpenPHError(rc, "vr2_device")
    return rows[:nnz.value], cols[:nnz.value], vals[:m.value], dims[:m.value]


def cubical_device(field, stream=None, device: int = 0):
    """Lower-star filtration of the cubical complex of an (s0, s1, s2)
    float32 CUDA tensor, built on the device. Returns (rows1, cols1, values,
    dims) as device tensors (the same matrix as complexes.cubical_lower_star)."""
    import torch

    s0, s1, s2 = field.shape
    m = ctypes.c_int64()
    nnz = ctypes.c_int64()
    rc = lib().openph_cubical_size(s0, s1, s2, ctypes.byref(m), ctypes.byref(nnz))
    if rc != 0:
        raise OpenPHError(rc, "cubical_device")
    dev = field.device
    rows = torch.empty(max(nnz.value, 1), dtype=torch.int32, device=dev)
    cols = torch.empty_like(rows)
    vals = torch.empty(max(m.value, 1), dtype=torch.float32, device=dev)
    dims = torch.empty(max(m.value, 1), dtype=torch.int32, device=dev)
    f = field.contiguous()
    rc = lib().openph_cubical_device(f.data_ptr(), s0, s1, s2, rows.data_ptr(), cols.data_ptr(),
                                     vals.data_ptr(), dims.data_ptr(), _stream_ptr(stream), device)
    if rc != 0:
        raise OpenPHError(rc, "cubical_device")
    return rows[:nnz.value], cols[:nnz.value], vals[:m.value], dims[:m.value]
