"""Served pms: Python handle over the openph_server C ABI (include/openph.h).

    with Server(workers=4) as srv:
        t = srv.submit(boundary)          # returns at once
        low0, ess, info = srv.wait(t)     # blocks until that request is done

ctypes releases the GIL during foreign calls, so submit() and wait() can be
called from different Python threads (an arrival loop and a collector).
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import Options, OpenPHError, lib

__all__ = ["Server", "ServerOptions", "RequestInfo", "Request"]

REQ_COO, REQ_VR, REQ_CUBICAL = 0, 1, 2
BORROW = 1


class ServerOptions(ctypes.Structure):
    _fields_ = [
        ("device", ctypes.c_int),
        ("workers", ctypes.c_int),
        ("worker_bytes", ctypes.c_int64),
        ("max_batch", ctypes.c_int),
        ("max_queue", ctypes.c_int),
        ("reduce", Options),
        ("stage_bytes", ctypes.c_int64),
    ]


class RequestInfo(ctypes.Structure):
    _fields_ = [
        ("submit_s", ctypes.c_double),
        ("start_s", ctypes.c_double),
        ("done_s", ctypes.c_double),
        ("batch_size", ctypes.c_int),
        ("worker", ctypes.c_int),
        ("num_iters", ctypes.c_int),
        ("status", ctypes.c_int),
        ("m", ctypes.c_int),
    ]

    @property
    def latency_s(self) -> float:
        return self.done_s - self.submit_s


class Request(ctypes.Structure):
    """openph_request: one request of any kind, optionally zero-copy."""
    _fields_ = [
        ("kind", ctypes.c_int),
        ("flags", ctypes.c_int),
        ("rows1", ctypes.c_void_p),
        ("cols1", ctypes.c_void_p),
        ("m", ctypes.c_int),
        ("nnz", ctypes.c_int),
        ("points", ctypes.c_void_p),
        ("n", ctypes.c_int),
        ("dim", ctypes.c_int),
        ("field", ctypes.c_void_p),
        ("s0", ctypes.c_int),
        ("s1", ctypes.c_int),
        ("s2", ctypes.c_int),
        ("low1", ctypes.c_void_p),
        ("ess", ctypes.c_void_p),
    ]


def _bind():
    L = lib()
    if getattr(L, "_server_bound", False):
        return L
    ip = ctypes.POINTER(ctypes.c_int)
    L.openph_server_create.restype = ctypes.c_int
    L.openph_server_create.argtypes = [ctypes.POINTER(ServerOptions), ctypes.POINTER(ctypes.c_void_p)]
    L.openph_server_submit.restype = ctypes.c_int
    L.openph_server_submit.argtypes = [ctypes.c_void_p, ip, ip, ctypes.c_int, ctypes.c_int,
                                       ctypes.POINTER(ctypes.c_int64)]
    L.openph_server_submit_vr.restype = ctypes.c_int
    L.openph_server_submit_vr.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int,
                                          ctypes.c_int, ctypes.POINTER(ctypes.c_int64)]
    L.openph_server_submit_cubical.restype = ctypes.c_int
    L.openph_server_submit_cubical.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int,
                                               ctypes.c_int, ctypes.c_int,
                                               ctypes.POINTER(ctypes.c_int64)]
    L.openph_server_submit_request.restype = ctypes.c_int
    L.openph_server_submit_request.argtypes = [ctypes.c_void_p, ctypes.POINTER(Request),
                                               ctypes.POINTER(ctypes.c_int64)]
    L.openph_server_wait.restype = ctypes.c_int
    L.openph_server_wait.argtypes = [ctypes.c_void_p, ctypes.c_int64, ip, ip,
                                     ctypes.POINTER(RequestInfo)]
    L.openph_server_stats.restype = ctypes.c_int
    L.openph_server_stats.argtypes = [ctypes.c_void_p] + [ctypes.POINTER(ctypes.c_int64)] * 3
    L.openph_server_destroy.restype = ctypes.c_int
    L.openph_server_destroy.argtypes = [ctypes.c_void_p]
    L._server_bound = True
    return L


class Server:
    def __init__(self, workers: int = 4, worker_bytes: int = 8 << 30, max_batch: int = 1024,
                 max_queue: int = 1 << 16, device: int = 0, reduce: Options | None = None,
                 stage_bytes: int = 0):
        L = _bind()
        o = ServerOptions()
        o.device = device
        o.workers = workers
        o.worker_bytes = worker_bytes
        o.max_batch = max_batch
        o.max_queue = max_queue
        o.stage_bytes = stage_bytes
        if reduce is not None:
            o.reduce = reduce
        h = ctypes.c_void_p()
        rc = L.openph_server_create(ctypes.byref(o), ctypes.byref(h))
        if rc != 0:
            raise OpenPHError(rc, "openph_server_create")
        self._h = h
        self._m = {}
        # borrowed requests: the arrays the worker reads and writes, kept
        # alive here until the ticket is waited for
        self._lent = {}

    def _submit_lent(self, q: Request, keep, m: int, what: str) -> int:
        low = np.empty(m, np.int32)
        ess = np.empty(m, np.int32)
        q.flags = BORROW
        q.low1 = low.ctypes.data
        q.ess = ess.ctypes.data
        t = ctypes.c_int64()
        rc = _bind().openph_server_submit_request(self._h, ctypes.byref(q), ctypes.byref(t))
        if rc != 0:
            raise OpenPHError(rc, what)
        self._m[t.value] = m
        self._lent[t.value] = (keep, low, ess)
        return t.value

    def submit_coo(self, rows1, cols1, m: int, borrow: bool = False) -> int:
        """A 1-based COO request. borrow=True submits without copying: the
        worker reads the arrays in place (they must not change until wait)
        and writes the results into arrays allocated here."""
        rows1 = np.ascontiguousarray(rows1, dtype=np.int32)
        cols1 = np.ascontiguousarray(cols1, dtype=np.int32)
        if borrow:
            q = Request(kind=REQ_COO, rows1=rows1.ctypes.data, cols1=cols1.ctypes.data, m=int(m),
                        nnz=int(rows1.shape[0]))
            return self._submit_lent(q, (rows1, cols1), int(m), "openph_server_submit_request")
        t = ctypes.c_int64()
        ip = ctypes.POINTER(ctypes.c_int)
        rc = _bind().openph_server_submit(self._h, rows1.ctypes.data_as(ip),
                                          cols1.ctypes.data_as(ip), int(m), int(rows1.shape[0]),
                                          ctypes.byref(t))
        if rc != 0:
            raise OpenPHError(rc, "openph_server_submit")
        self._m[t.value] = int(m)
        return t.value

    def submit(self, boundary) -> int:
        r, c = boundary.coo1()
        return self.submit_coo(r, c, boundary.m)

    def submit_points(self, points, borrow: bool = False) -> int:
        """A VR 2-skeleton request: (n, d) points, d <= 7; the worker builds
        the matrix on the device (openph_vr2_device's order)."""
        p = np.ascontiguousarray(points, dtype=np.float64)
        n, d = p.shape
        if borrow:
            q = Request(kind=REQ_VR, points=p.ctypes.data, n=n, dim=d)
            m = n + n * (n - 1) // 2 + n * (n - 1) * (n - 2) // 6
            return self._submit_lent(q, p, m, "openph_server_submit_request")
        t = ctypes.c_int64()
        rc = _bind().openph_server_submit_vr(self._h, p.ctypes.data, n, d, ctypes.byref(t))
        if rc != 0:
            raise OpenPHError(rc, "openph_server_submit_vr")
        self._m[t.value] = n + n * (n - 1) // 2 + n * (n - 1) * (n - 2) // 6
        return t.value

    def submit_field(self, field, borrow: bool = False) -> int:
        """A cubical lower-star request: an (s0, s1, s2) float32 grid; the
        worker builds the matrix on the device (openph_cubical_device's
        order)."""
        f = np.ascontiguousarray(field, dtype=np.float32)
        s0, s1, s2 = f.shape
        if borrow:
            q = Request(kind=REQ_CUBICAL, field=f.ctypes.data, s0=s0, s1=s1, s2=s2)
            m = (2 * s0 - 1) * (2 * s1 - 1) * (2 * s2 - 1)
            return self._submit_lent(q, f, m, "openph_server_submit_request")
        t = ctypes.c_int64()
        rc = _bind().openph_server_submit_cubical(self._h, f.ctypes.data, s0, s1, s2,
                                                  ctypes.byref(t))
        if rc != 0:
            raise OpenPHError(rc, "openph_server_submit_cubical")
        self._m[t.value] = (2 * s0 - 1) * (2 * s1 - 1) * (2 * s2 - 1)
        return t.value

    def wait(self, ticket: int):
        """Returns (low0, ess, RequestInfo); low0 is 0-based (-1 = zero)."""
        m = self._m.pop(ticket)
        lent = self._lent.pop(ticket, None)
        if lent is not None:
            _, low, ess = lent  # the worker wrote the results there
            lp = ep = None
        else:
            low = np.empty(m, np.int32)
            ess = np.empty(m, np.int32)
            ip = ctypes.POINTER(ctypes.c_int)
            lp, ep = low.ctypes.data_as(ip), ess.ctypes.data_as(ip)
        info = RequestInfo()
        rc = _bind().openph_server_wait(self._h, ticket, lp, ep, ctypes.byref(info))
        if rc != 0:
            raise OpenPHError(rc, "openph_server_wait")
        low -= 1
        return low, ess, info

    def stats(self):
        q, b, c = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
        _bind().openph_server_stats(self._h, ctypes.byref(q), ctypes.byref(b), ctypes.byref(c))
        return {"queued": q.value, "batches": b.value, "completed": c.value}

    def close(self):
        if self._h:
            _bind().openph_server_destroy(self._h)  # finishes what is queued
            self._h = None
            self._lent.clear()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
