"""Served pms on the GPU: results in a concurrent mix equal each request
reduced alone (and the golden fixtures / oracle), from several submitting
threads, with requests of very different sizes."""
import json
import os
import sys
import threading
import time

import numpy as np
import pytest

import oracle
import openph_amd
from openph_amd import complexes as C

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(__file__), "golden")
INDEX = json.load(open(os.path.join(GOLD, "index.json")))
sys.path.insert(0, os.path.dirname(__file__))
from test_oracle import _gen  # noqa: E402


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    openph_amd.lib()


def test_server_mix_equals_alone():
    from openph_amd.server import Server
    print("server mix", flush=True)
    names = sorted(INDEX)
    mats = [_gen(n) for n in names] + C.batched_mix(12, seed=11) + [C.cubical64(seed=3, side=20)]
    alone = [openph_amd.reduce(b, "pms")[:2] for b in mats]
    with Server(workers=3, worker_bytes=256 << 20) as srv:
        tickets = [None] * len(mats)

        def submit(idx):
            for i in idx:
                tickets[i] = srv.submit(mats[i])

        th = [threading.Thread(target=submit, args=(list(range(k, len(mats), 3)),))
              for k in range(3)]
        for t in th:
            t.start()
        for t in th:
            t.join()
        infos = []
        for i, b in enumerate(mats):
            low, ess, info = srv.wait(tickets[i])
            infos.append(info)
            assert info.status == 0 and info.m == b.m
            assert np.array_equal(low, alone[i][0]), b.name
            assert np.array_equal(ess, alone[i][1]), b.name
            if i < len(names):
                f = np.load(os.path.join(GOLD, names[i] + ".npz"))
                assert np.array_equal(low, f["low"])
        st = srv.stats()
    assert st["completed"] == len(mats)
    assert all(i.done_s >= i.start_s >= i.submit_s for i in infos)
    # the requests were batched: fewer batches than requests
    assert st["batches"] < len(mats)


def test_server_small_budget_runs_oversized_request_alone():
    from openph_amd.server import Server
    print("server budget", flush=True)
    big = C.torus100()
    small = [_gen("vr12"), _gen("random1")]
    with Server(workers=1, worker_bytes=1 << 20) as srv:
        tb = srv.submit(big)
        ts = [srv.submit(b) for b in small]
        low, ess, info = srv.wait(tb)
        f = np.load(os.path.join(GOLD, "torus100.npz"))
        assert np.array_equal(low, f["low"]) and info.batch_size == 1
        for b, t in zip(small, ts):
            low, ess, _ = srv.wait(t)
            assert np.array_equal(low, oracle.pms(b).low)


def test_server_rejects_malformed_request_without_poisoning_others():
    from openph_amd.server import Server
    print("server malformed", flush=True)
    r = np.array([1, 2, 1, 3, 2, 3, 4, 5, 6], np.int32)
    c = np.array([4, 4, 5, 5, 6, 6, 7, 7, 7], np.int32)
    bad_r = r.copy()
    bad_r[0] = 5  # row >= col
    with Server(workers=1) as srv:
        tg0 = srv.submit_coo(r, c, 7)
        with pytest.raises(openph_amd.OpenPHError) as e:
            srv.submit_coo(bad_r, c, 7)
        assert e.value.status == -1
        bad_c = c.copy()
        bad_c[2], bad_c[6] = bad_c[6], bad_c[2]
        with pytest.raises(openph_amd.OpenPHError) as e:
            srv.submit_coo(r, bad_c, 7)
        assert e.value.status == -2
        tg = srv.submit_coo(r, c, 7)
        for t in (tg0, tg):
            low, ess, info = srv.wait(t)
            assert (low + 1).tolist() == [0, 0, 0, 2, 3, 0, 6]


def test_server_builds_point_clouds_and_fields_on_the_device():
    """Requests given as point clouds and grids, mixed with COO requests in
    the same batches: each result equals the host-built matrix reduced
    alone."""
    from openph_amd.server import Server
    print("server built requests", flush=True)
    rng = np.random.default_rng(11)
    items = []
    for k in range(10):
        if k % 3 == 2:
            f = C.smoothed_field(rng, int(rng.integers(6, 14)))
            items.append(("field", f, C.cubical_lower_star(f)))
        else:
            p = C.points_cube(rng, int(rng.integers(5, 30)))
            items.append(("points" if k % 3 == 0 else "coo", p, C.vr_2skeleton(p)))
    with Server(workers=2) as srv:
        tickets = []
        for kind, x, b in items:
            if kind == "field":
                tickets.append(srv.submit_field(x))
            elif kind == "points":
                tickets.append(srv.submit_points(x))
            else:
                tickets.append(srv.submit(b))
        for t, (kind, x, b) in zip(tickets, items):
            low, ess, info = srv.wait(t)
            ref = oracle.pms(b)
            assert np.array_equal(low, ref.low) and np.array_equal(ess, ref.ess), kind
            assert info.m == b.m
        with pytest.raises(openph_amd.OpenPHError):
            srv.submit_points(np.full((4, 3), np.nan))


def test_server_admits_into_a_running_batch_and_releases_members_early():
    """One worker: a large cubical grid starts a batch; a small grid of the
    same class submitted while it runs joins that batch at an iteration
    boundary and leaves it as soon as its own lows stop repeating, before
    the large one is done. Both equal the oracle."""
    from openph_amd.server import Server
    print("server continuous batching", flush=True)
    big = C.cubical_lower_star(C.smoothed_field(np.random.default_rng(21), 80))
    small = C.cubical_lower_star(C.smoothed_field(np.random.default_rng(22), 10))
    with Server(workers=1) as srv:
        rb, cb = big.coo1()
        tb = srv.submit_coo(rb, cb, big.m)
        t0 = time.perf_counter()
        while srv.stats()["queued"] and time.perf_counter() - t0 < 30:
            time.sleep(0.0005)
        ts = srv.submit(small)
        low_s, ess_s, info_s = srv.wait(ts)
        low_b, ess_b, info_b = srv.wait(tb)
    ref_s, ref_b = oracle.pms(small), oracle.pms(big)
    assert np.array_equal(low_s, ref_s.low) and np.array_equal(ess_s, ref_s.ess)
    assert np.array_equal(low_b, ref_b.low) and np.array_equal(ess_b, ref_b.ess)
    assert info_s.worker == info_b.worker == 0
    assert info_s.start_s > info_b.start_s  # joined the running batch later
    assert info_s.done_s < info_b.done_s    # and left it first
    assert info_b.batch_size >= 2


def _full_simplex(n):
    """All faces of the simplex on n vertices (by dimension, then
    lexicographic): column lengths 0, 2, ..., n."""
    import itertools
    cells = [c for k in range(1, n + 1) for c in itertools.combinations(range(n), k)]
    pos = {c: i for i, c in enumerate(cells)}
    rows, cols = [], []
    for j, c in enumerate(cells):
        if len(c) > 1:
            f = sorted(pos[c[:i] + c[i + 1:]] for i in range(len(c)))
            rows += [x + 1 for x in f]
            cols += [j + 1] * len(f)
    return np.array(rows, np.int32), np.array(cols, np.int32), len(cells)


def _cube_cells(d):
    """All faces of the d-cube (coordinates 0 and 2 for the ends, 1 for the
    interval; by dimension, then index): column lengths 0, 2, 4, ..., 2d."""
    import itertools
    cells = sorted(itertools.product((0, 1, 2), repeat=d), key=lambda c: (sum(x == 1 for x in c), c))
    pos = {c: i for i, c in enumerate(cells)}
    rows, cols = [], []
    for j, c in enumerate(cells):
        f = sorted(pos[c[:i] + (e,) + c[i + 1:]] for i, x in enumerate(c) if x == 1 for e in (0, 2))
        rows += [x + 1 for x in f]
        cols += [j + 1] * len(f)
    return np.array(rows, np.int32), np.array(cols, np.int32), len(cells)


def test_server_keeps_batches_within_the_pseudo_dimension_limit():
    """The full simplex on 16 vertices (column lengths 0, 2..16) and the
    faces of a 9-cube (0, 2, 4, ..., 18): 17 distinct lengths over both,
    more than a batch can tell apart (OPENPH_MAX_DIMS = 16). They never
    share a batch, and both succeed. Both are contractible: one essential
    class each."""
    from openph_amd.server import Server
    print("server length union", flush=True)
    cases = [_full_simplex(16), _cube_cells(9)]
    with Server(workers=1) as srv:
        tickets = [srv.submit_coo(*x) for x in cases]
        res = [srv.wait(t) for t in tickets]
    for (r, c, m), (low, ess, info) in zip(cases, res):
        assert info.status == 0 and info.batch_size == 1
        ref = openph_amd.ph("pms", r, c, m)
        assert np.array_equal(low + 1, ref[0]) and np.array_equal(ess, ref[1])
        assert ess.sum() == 1


def test_server_ticket_is_waited_for_once():
    """A second wait on a ticket that another thread is waiting for is
    refused (OPENPH_EINVAL) instead of reading a request that the first
    waiter frees."""
    from openph_amd.server import Server
    import ctypes
    print("server double wait", flush=True)
    b = C.cubical_lower_star(C.smoothed_field(np.random.default_rng(5), 48))
    with Server(workers=1) as srv:
        t = srv.submit(b)
        out = {}
        th = threading.Thread(target=lambda: out.setdefault("r", srv.wait(t)))
        th.start()
        time.sleep(0.002)
        low = np.empty(b.m, np.int32)
        ess = np.empty(b.m, np.int32)
        ip = ctypes.POINTER(ctypes.c_int)
        from openph_amd.server import _bind
        rc = _bind().openph_server_wait(srv._h, t, low.ctypes.data_as(ip), ess.ctypes.data_as(ip),
                                        None)
        th.join()
    assert rc == -1
    assert np.array_equal(out["r"][0], oracle.pms(b).low)


def test_server_borrowed_requests_equal_copied_ones():
    """Zero-copy requests (openph_server_submit_request with OPENPH_BORROW
    and caller-owned outputs) of all three kinds, mixed with copied ones in
    the same batches and, under a small budget, one large enough to run
    alone: each equals the oracle."""
    from openph_amd.server import Server
    print("server borrowed requests", flush=True)
    rng = np.random.default_rng(31)
    items = []
    for k in range(9):
        if k % 3 == 2:
            f = C.smoothed_field(rng, int(rng.integers(6, 14)))
            items.append(("field", f, C.cubical_lower_star(f)))
        else:
            p = C.points_cube(rng, int(rng.integers(5, 30)))
            items.append(("points" if k % 3 == 0 else "coo", p, C.vr_2skeleton(p)))
    big = C.cubical_lower_star(C.smoothed_field(np.random.default_rng(32), 40))
    for budget in (8 << 30, 64 << 20):
        with Server(workers=2, worker_bytes=budget) as srv:
            tickets = []
            for k, (kind, x, b) in enumerate(items):
                lend = k % 2 == 0
                if kind == "field":
                    tickets.append(srv.submit_field(x, borrow=lend))
                elif kind == "points":
                    tickets.append(srv.submit_points(x, borrow=lend))
                else:
                    r, c = b.coo1()
                    tickets.append(srv.submit_coo(r, c, b.m, borrow=lend))
            rb, cb = big.coo1()
            tbig = srv.submit_coo(rb, cb, big.m, borrow=True)
            # a lent COO is validated after submit: its error comes at wait,
            # and the other requests are not affected
            bad = rb.copy()
            j = int(np.flatnonzero(cb[1:] == cb[:-1])[0])  # two entries of one column
            bad[j], bad[j + 1] = bad[j + 1], bad[j]          # rows no longer ascending
            tbad = srv.submit_coo(bad, cb, big.m, borrow=True)
            with pytest.raises(openph_amd.OpenPHError) as e:
                srv.wait(tbad)
            assert e.value.status == -2  # OPENPH_EUNSORTED
            for t, (kind, x, b) in zip(tickets, items):
                low, ess, info = srv.wait(t)
                ref = oracle.pms(b)
                assert info.status == 0 and info.m == b.m
                assert np.array_equal(low, ref.low) and np.array_equal(ess, ref.ess), kind
            low, ess, info = srv.wait(tbig)
            ref = oracle.pms(big)
            assert np.array_equal(low, ref.low) and np.array_equal(ess, ref.ess)
            if budget < (1 << 30):
                assert info.batch_size == 1


def test_server_full_queue_refuses_and_keeps_serving():
    """A queue of two: submits beyond it are refused with OPENPH_EBUSY (copied
    and lent alike, with a small pinned arena that the refused requests must
    not keep), and every accepted request still equals the oracle."""
    from openph_amd.server import Server
    print("server full queue", flush=True)
    b = C.cubical_lower_star(C.smoothed_field(np.random.default_rng(41), 24))
    r, c = b.coo1()
    ref = oracle.pms(b)
    busy = 0
    with Server(workers=1, max_queue=2, stage_bytes=6 << 20) as srv:
        for lend in (False, True):
            tickets = []
            for _ in range(12):
                try:
                    tickets.append(srv.submit_coo(r, c, b.m, borrow=lend))
                except openph_amd.OpenPHError as e:
                    assert e.status == -8  # OPENPH_EBUSY
                    busy += 1
            for t in tickets:
                low, ess, info = srv.wait(t)
                assert info.status == 0
                assert np.array_equal(low, ref.low) and np.array_equal(ess, ref.ess)
    assert busy > 0
