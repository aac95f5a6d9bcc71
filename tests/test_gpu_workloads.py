"""GPU parity on the two full-size workloads whose outputs are pinned by
digests (tests/golden/make_digests.py): the 512-filtration batched mix and a
served request stream. Every member's low and ess equal the oracle's
(digests of arrays the oracle's pms and twist agree on), and in the served
stream also the same matrix reduced alone, as src/test/test_ph.m:51-66
compares each algorithm with a reference reduction."""
import hashlib
import json
import os
import threading
import time

import numpy as np
import pytest

import openph_amd
from openph_amd import complexes as C

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def say(*a):
    print(*a, flush=True)


def digest(a) -> str:
    return hashlib.sha256(np.ascontiguousarray(a, dtype="<i4").tobytes()).hexdigest()[:16]


def coo_digest(b) -> str:
    r, c = b.coo1()
    h = hashlib.sha256()
    h.update(np.int64(b.m).tobytes())
    h.update(np.ascontiguousarray(r, dtype="<i4").tobytes())
    h.update(np.ascontiguousarray(c, dtype="<i4").tobytes())
    return h.hexdigest()[:16]


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    openph_amd.lib()


@pytest.fixture(scope="module")
def batch512():
    import torch
    gold = json.load(open(os.path.join(GOLD, "batched512.json")))["members"]
    mats = C.batched_mix(512, seed=0)
    assert [coo_digest(b) for b in mats] == [g["coo"] for g in gold]
    ms = np.array([b.m for b in mats], np.int32)
    nn = np.array([b.nnz for b in mats], np.int32)
    rr, cc = zip(*(b.coo1() for b in mats))
    dev = torch.device("cuda:0")
    rows = torch.from_numpy(np.concatenate(rr)).to(dev)
    cols = torch.from_numpy(np.concatenate(cc)).to(dev)
    return gold, ms, nn, rows, cols


def _check_batch(gold, ms, low, ess):
    low = low.cpu().numpy()
    ess = ess.cpu().numpy()
    o = 0
    bad = []
    for k, (g, m) in enumerate(zip(gold, ms)):
        if digest(low[o:o + m] - 1) != g["low"] or digest(ess[o:o + m]) != g["ess"]:
            bad.append(k)
        o += m
    assert not bad, f"members differing from the oracle: {bad[:10]}"


@pytest.mark.parametrize("tuned", [False, True])
def test_batched512_every_member_equals_the_oracle(batch512, tuned):
    import torch
    gold, ms, nn, rows, cols = batch512
    say("batched512", "tuned" if tuned else "default options")
    M = int(ms.sum())
    low = torch.full((M,), -7, dtype=torch.int32, device="cuda")
    ess = torch.full((M,), -7, dtype=torch.int32, device="cuda")
    opt = None
    if tuned:
        opt, rows_ = openph_amd.autotune_batch_device(ms, nn, rows, cols, low, ess, reps=1)
        assert all(r["same"] == 1 for r in rows_)
        low.fill_(-7)
        ess.fill_(-7)
    st = openph_amd.pms_batch_device(ms, nn, rows, cols, low, ess, options=opt)
    torch.cuda.synchronize()
    say("batched512 stats", st.num_iters, st.additions, st.lanes, st.group_lds, st.schedule)
    _check_batch(gold, ms, low, ess)
    # the oracle's additions over the 512 reductions, one block-diagonal matrix
    assert st.additions == 229510826


def test_served_stream_equals_oracle_and_alone():
    """17 requests of the serving mix (VR N up to 190 as COO, cubical grids
    of side up to 85 as fields built on the device) submitted by a thread
    with Poisson arrivals (mean gap 4 ms) to a 4-worker server."""
    import torch

    from openph_amd.server import Server
    gold = json.load(open(os.path.join(GOLD, "served17.json")))["requests"]
    rng = np.random.default_rng(0)
    reqs = []
    for g in gold:
        kind, x = C.served_input(rng)
        assert kind == g["kind"] and x.shape[0] == g["size"]
        if kind == "points":
            b = C.vr_2skeleton(x)
            assert coo_digest(b) == g["coo"]
            reqs.append(("coo", b))
        else:
            reqs.append(("field", x))
    say("served17: requests generated")
    # each request reduced alone
    alone = []
    for kind, x in reqs:
        if kind == "coo":
            low, ess, _ = openph_amd.reduce(x, "pms")
        else:
            rows, cols, vals, _ = openph_amd.cubical_device(torch.from_numpy(x).cuda())
            m = vals.numel()
            lo = torch.empty(m, dtype=torch.int32, device="cuda")
            es = torch.empty_like(lo)
            openph_amd.pms_device(rows, cols, m, lo, es)
            torch.cuda.synchronize()
            low, ess = lo.cpu().numpy() - 1, es.cpu().numpy()
        alone.append((low, ess))
    for (low, ess), g in zip(alone, gold):
        assert digest(low) == g["low"] and digest(ess) == g["ess"]
    say("served17: alone equal the oracle")
    arrivals = np.cumsum(np.random.default_rng(1).exponential(0.004, len(reqs)))
    tickets = [None] * len(reqs)
    with Server(workers=4) as srv:
        def submitter():
            t0 = time.perf_counter()
            for k, (kind, x) in enumerate(reqs):
                dt = arrivals[k] - (time.perf_counter() - t0)
                if dt > 0:
                    time.sleep(dt)
                if kind == "coo":
                    r, c = x.coo1()
                    tickets[k] = srv.submit_coo(r, c, x.m)
                else:
                    tickets[k] = srv.submit_field(x)

        th = threading.Thread(target=submitter)
        th.start()
        results = []
        for k in range(len(reqs)):
            while tickets[k] is None:
                time.sleep(0.0005)
            results.append(srv.wait(tickets[k]))
        th.join()
    for k, ((low, ess, info), g) in enumerate(zip(results, gold)):
        assert info.status == 0 and info.m == g["m"], k
        assert digest(low) == g["low"] and digest(ess) == g["ess"], k
        assert np.array_equal(low, alone[k][0]) and np.array_equal(ess, alone[k][1]), k
    say("served17: served equal alone and the oracle")
