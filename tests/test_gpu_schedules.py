"""GPU tests of the loop schedules and of the drop-in's trackers.

Schedules: the host either reads the device state twice per iteration
(sync), or queues chunks of iterations whose kernels stop themselves on the
device (queued), or replays such chunks from a cached hipGraph (graph). All
three run the same iterations, so low/ess and the iteration, addition and
merged-element counts must equal the oracle's snapshot schedule in each,
including across pool compactions, errors and edge cases.

Trackers: the four err_* arrays of the reference's track()
(src/cuda/algorithms/tracker.cu:92-134) per iteration, scaled as
src/cuda/openph.cu:188-193 does. The GPU runs the oracle's snapshot
schedule, so its per-iteration integers must equal the oracle's and the
floats made from them must be identical.
"""
import ctypes
import json
import os
import sys

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

O = openph_amd.Options


def say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    openph_amd.lib()


def _opt(schedule, chunk=0, **kw):
    o = O()
    o.schedule = schedule
    o.chunk_iters = chunk
    for k, v in kw.items():
        setattr(o, k, v)
    return o


@pytest.mark.parametrize("schedule,chunk", [(O.SYNC, 0), (O.QUEUED, 1), (O.QUEUED, 2),
                                            (O.QUEUED, 0), (O.GRAPH, 1), (O.GRAPH, 3)])
@pytest.mark.parametrize("name", ["vr12", "random3", "vr40", "torus100", "cubical16"])
def test_schedules_match_golden_and_counts(name, schedule, chunk):
    say("schedule", name, schedule, chunk)
    f = np.load(os.path.join(GOLD, name + ".npz"))
    b = _gen(name)
    snap = INDEX[name]["pms_snapshot"]
    for rep in range(2):  # the second run starts with a chunk of the learnt length
        low, ess, st = openph_amd.reduce(b, "pms", _opt(schedule, chunk))
        assert np.array_equal(low, f["low"]) and np.array_equal(ess, f["ess"]), rep
        assert st.schedule == schedule or (schedule == O.GRAPH and st.schedule == O.QUEUED)
        assert st.num_iters == snap["num_iters"]
        assert st.additions == snap["additions"]
        assert st.merge_elems == snap["merge_elems"]
        assert st.phase2_launches == snap["num_iters"] - 1
        assert st.host_reads >= 1


def test_graph_schedule_is_used():
    """The graph schedule really replays a graph (it falls back to queued
    launches only when capture is refused, which this runtime should not do)."""
    b = _gen("torus100")
    _, _, st = openph_amd.reduce(b, "pms", _opt(O.GRAPH))
    assert st.schedule == O.GRAPH
    # a repeated reduction of the same size reads the control block once
    _, _, st = openph_amd.reduce(b, "pms", _opt(O.GRAPH))
    assert st.host_reads == 1


def test_graph_replays_interleaved_with_other_work():
    """One cached graph replayed several times on device buffers, with other
    reductions (other sizes, other schedules) and eager torch work between
    the replays: every replay equals the eager (sync) result."""
    import torch
    say("graph replays")
    dev = torch.device("cuda:0")
    b = _gen("torus100")
    f = np.load(os.path.join(GOLD, "torus100.npz"))
    snap = INDEX["torus100"]["pms_snapshot"]
    r, c = b.coo1()
    R = torch.from_numpy(r).to(dev)
    Cc = torch.from_numpy(c).to(dev)
    low = torch.empty(b.m, dtype=torch.int32, device=dev)
    ess = torch.empty_like(low)
    st = openph_amd.pms_device(R, Cc, b.m, low, ess, options=_opt(O.SYNC))
    torch.cuda.synchronize()
    eager_low, eager_ess = low.cpu().numpy().copy(), ess.cpu().numpy().copy()
    assert np.array_equal(eager_low, f["low"] + 1) and np.array_equal(eager_ess, f["ess"])
    other = C.vr40()
    ro, co = other.coo1()
    fo = np.load(os.path.join(GOLD, "vr40.npz"))
    for k in range(4):
        low.fill_(-7)
        ess.fill_(-7)
        st = openph_amd.pms_device(R, Cc, b.m, low, ess, options=_opt(O.GRAPH))
        x = torch.arange(1 << 20, device=dev, dtype=torch.float32).sum()  # eager work
        torch.cuda.synchronize()
        assert float(x) > 0
        assert st.schedule == O.GRAPH
        assert st.num_iters == snap["num_iters"] and st.additions == snap["additions"]
        assert np.array_equal(low.cpu().numpy(), eager_low), k
        assert np.array_equal(ess.cpu().numpy(), eager_ess), k
        # another size on the same workspace in between
        lo2, es2, _ = openph_amd.reduce(other, "pms", _opt(O.QUEUED if k % 2 else O.GRAPH))
        assert np.array_equal(lo2, fo["low"]) and np.array_equal(es2, fo["ess"])


@pytest.mark.parametrize("schedule", [O.SYNC, O.QUEUED, O.GRAPH])
def test_schedules_compact_and_stay_exact(schedule):
    say("small pool", schedule)
    b = C.torus100()
    f = np.load(os.path.join(GOLD, "torus100.npz"))
    opt = _opt(schedule, 2, slot_elems=1, pool_elems=b.nnz + b.m + 4096)
    low, ess, st = openph_amd.reduce(b, "pms", opt)
    assert st.compactions > 0 and st.overflows > 0
    assert np.array_equal(low, f["low"]) and np.array_equal(ess, f["ess"])


@pytest.mark.parametrize("schedule", [O.SYNC, O.QUEUED, O.GRAPH])
def test_schedules_pool_budget_error(schedule):
    say("pool budget", schedule)
    b = C.torus100()
    f = np.load(os.path.join(GOLD, "torus100.npz"))
    opt = _opt(schedule, 2, pool_elems=b.nnz, pool_max_elems=b.nnz)
    try:
        low, ess, _ = openph_amd.reduce(b, "pms", opt)
    except openph_amd.OpenPHError as e:
        assert e.status == -4
        return
    assert np.array_equal(low, f["low"]) and np.array_equal(ess, f["ess"])


@pytest.mark.parametrize("schedule", [O.SYNC, O.QUEUED, O.GRAPH])
def test_schedules_iteration_cap(schedule):
    say("iteration cap", schedule)
    b = C.torus100()
    with pytest.raises(openph_amd.OpenPHError) as e:
        openph_amd.reduce(b, "pms", _opt(schedule, 2, max_iters=2))
    assert e.value.status == -6
    f = np.load(os.path.join(GOLD, "torus100.npz"))
    low, _, _ = openph_amd.reduce(b, "pms", _opt(schedule, 2))
    assert np.array_equal(low, f["low"])


@pytest.mark.parametrize("schedule", [O.SYNC, O.QUEUED, O.GRAPH])
def test_schedules_edge_cases(schedule):
    say("edge cases", schedule)
    opt = _opt(schedule)
    b = C.Boundary(colptr=np.zeros(6, np.int64), rows=np.zeros(0, np.int32),
                   dim=np.zeros(5, np.int32), value=np.zeros(5))
    low, ess, st = openph_amd.reduce(b, "pms", opt)
    assert (low == -1).all() and (ess == 1).all() and st.num_iters == 2
    # a single edge: one pair found by Phase 0, converged after iteration 1
    b = C.Boundary(colptr=np.array([0, 0, 0, 2], np.int64), rows=np.array([0, 1], np.int32),
                   dim=np.array([0, 0, 1], np.int32), value=np.zeros(3))
    low, ess, st = openph_amd.reduce(b, "pms", opt)
    assert low.tolist() == [-1, -1, 1] and ess.tolist() == [1, 0, 0]


@pytest.mark.parametrize("schedule", [O.QUEUED, O.GRAPH])
def test_schedules_batch_matches_plain(schedule):
    say("batch", schedule)
    mats = C.batched_mix(24, seed=5)
    out, st = openph_amd.pms_batch(mats, _opt(schedule))
    for b, (low, ess) in zip(mats, out):
        ref = oracle.pms(b)
        assert np.array_equal(low, ref.low) and np.array_equal(ess, ref.ess)


def _expected_trackers(trk, m, lt1, n):
    """The reference's floats from the oracle's per-iteration integers."""
    f32 = np.float32
    norm1 = f32(np.abs(lt1.astype(np.int64)).sum())
    norminf = f32(np.abs(lt1).max())
    paired = np.zeros(m, bool)
    j = np.flatnonzero(lt1 > 0)
    paired[j] = True
    paired[lt1[j] - 1] = True
    ess_true = int((~paired).sum())
    lone = np.full(m + 1, -1, f32)
    linf = np.full(m + 1, -1, f32)
    redu = np.full(m + 1, -1, f32)
    essr = np.full(m + 1, -1, f32)
    lone[:n] = trk[:, 0].astype(f32)
    linf[:n] = trk[:, 1].astype(f32)
    redu[:n] = trk[:, 2].astype(f32) / f32(m)
    hat = trk[:, 3]
    essr[:n] = np.where(hat == ess_true, f32(1), f32(ess_true) / hat.astype(f32))
    return linf / norminf, lone / norm1, redu, essr


@pytest.mark.parametrize("schedule", [0, 1, 2, 3])
@pytest.mark.parametrize("name", ["vr12", "random1", "vr40", "torus100", "cubical16"])
def test_trackers_match_oracle(name, schedule):
    """Every schedule fills the same trackers: the sync one reads them after
    each iteration, the queued and graph ones keep them per iteration on the
    device and read them once per chunk (the default)."""
    say("trackers", name, "schedule", schedule)
    f = np.load(os.path.join(GOLD, name + ".npz"))
    b = _gen(name)
    r, c = b.coo1()
    lt1 = (f["low"] + 1).astype(np.int32)
    o = openph_amd.ph("pms", r, c, b.m, 8, lt1, schedule=schedule)
    low, ess, err_linf, err_lone, err_redu, err_ess, time_track, nit = o
    n = int(nit[0])
    ref = oracle.pms(b, snapshot=True, low_true=f["low"].astype(np.int32))
    assert n == ref.num_iters
    assert np.array_equal(low, lt1) and np.array_equal(ess, f["ess"])
    e_linf, e_lone, e_redu, e_ess = _expected_trackers(ref.trk, b.m, lt1, n)
    assert np.array_equal(err_linf, e_linf)
    assert np.array_equal(err_lone, e_lone)
    assert np.array_equal(err_redu, e_redu)
    assert np.array_equal(err_ess, e_ess)
    assert time_track[0] == 0 and (time_track[1:n] > 0).all() and (time_track[n:] == 0).all()
    # the last tracked state is the reduced matrix
    assert err_lone[n - 1] == 0 and err_linf[n - 1] == 0 and err_redu[n - 1] == 0
    assert err_ess[n - 1] == 1


@pytest.mark.parametrize("schedule", [1, 2, 3])
def test_trackers_with_suspended_passes(schedule):
    """With a pool too small for Phase II, passes are suspended and finished
    after a compaction; the trackers and iteration times still describe whole
    iterations, equal to the oracle's."""
    say("trackers, small pool", schedule)
    f = np.load(os.path.join(GOLD, "torus100.npz"))
    b = _gen("torus100")
    r, c = b.coo1()
    lt1 = (f["low"] + 1).astype(np.int32)
    base = openph_amd.Options()
    base.slot_elems = 1
    base.pool_elems = b.nnz + b.m + 4096
    o = openph_amd.ph("pms", r, c, b.m, 8, lt1, schedule=schedule, options=base)
    low, ess, err_linf, err_lone, err_redu, err_ess, time_track, nit = o
    n = int(nit[0])
    ref = oracle.pms(b, snapshot=True, low_true=f["low"].astype(np.int32))
    assert n == ref.num_iters
    assert np.array_equal(low, lt1) and np.array_equal(ess, f["ess"])
    e_linf, e_lone, e_redu, e_ess = _expected_trackers(ref.trk, b.m, lt1, n)
    assert np.array_equal(err_linf, e_linf)
    assert np.array_equal(err_lone, e_lone)
    assert np.array_equal(err_redu, e_redu)
    assert np.array_equal(err_ess, e_ess)
    assert time_track[0] == 0 and (time_track[1:n] > 0).all() and (time_track[n:] == 0).all()


def test_dropin_void_openph_fills_trackers_like_ph():
    """openph() itself (the C symbol the mex gateway calls) with outputs
    initialised as src/matlab/ph/ph.cu:72-78 does."""
    say("void openph trackers")
    f = np.load(os.path.join(GOLD, "vr40.npz"))
    b = _gen("vr40")
    r, c = b.coo1()
    m = b.m
    lt1 = (f["low"] + 1).astype(np.int32)
    low = np.full(m, -1, np.int32)
    ess = np.full(m, -1, np.int32)
    errs = [np.full(m + 1, -1, np.float32) for _ in range(4)]
    tt = np.zeros(m + 1, np.float32)
    nit = np.zeros(1, np.int32)
    ip = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))  # noqa: E731
    fp = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))  # noqa: E731
    openph_amd.lib().openph(b"pms", ip(r), ip(c), m, 8, ip(lt1), r.shape[0], ip(low), ip(ess),
                            fp(errs[0]), fp(errs[1]), fp(errs[2]), fp(errs[3]), fp(tt), ip(nit), 0)
    o = openph_amd.ph("pms", r, c, m, 8, lt1)
    assert nit[0] == o[7][0] > 1
    assert np.array_equal(low, o[0]) and np.array_equal(ess, o[1])
    for k in range(4):
        assert np.array_equal(errs[k], o[2 + k]), k
    # an error leaves low as initialised and reports the status in num_iters
    bad_r = r.copy()
    bad_r[0] = m + 5
    low[:] = -1
    openph_amd.lib().openph(b"pms", ip(bad_r), ip(c), m, 8, ip(lt1), r.shape[0], ip(low),
                            ip(ess), fp(errs[0]), fp(errs[1]), fp(errs[2]), fp(errs[3]), fp(tt),
                            ip(nit), 0)
    assert nit[0] < 0 and (low == -1).all()
    openph_amd.lib().openph(b"no_such_algo", ip(r), ip(c), m, 8, ip(lt1), r.shape[0], ip(low),
                            ip(ess), fp(errs[0]), fp(errs[1]), fp(errs[2]), fp(errs[3]), fp(tt),
                            ip(nit), 0)
    assert nit[0] == -3


def test_dropin_with_trackers_does_not_force_the_sync_schedule():
    """The reference's callers always pass low_true (zeros,
    numerics/benchmark_datasets.m:53): with it the drop-in still queues its
    iterations (one control-block read per chunk) and its outputs and
    trackers equal the sync schedule's."""
    say("dropin schedule")
    f = np.load(os.path.join(GOLD, "torus100.npz"))
    b = _gen("torus100")
    r, c = b.coo1()
    m = b.m
    lt = np.zeros(m, np.int32)
    outs = {}
    for sched in (1, 0, 0):
        low = np.full(m, -1, np.int32)
        ess = np.full(m, -1, np.int32)
        trk = np.full((5, m + 1), -1.0, np.float32)
        trk[4, :] = 0.0
        opt = openph_amd.Options()
        opt.track = 1
        opt.schedule = sched
        st = openph_amd.Stats()
        ip = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))  # noqa: E731
        rc = openph_amd.lib().openph_v2(b"pms", ip(r), ip(c), m, r.shape[0], ip(lt),
                                        ctypes.byref(opt), ip(low), ip(ess),
                                        trk.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                                        ctypes.byref(st))
        assert rc == 0
        assert np.array_equal(low, f["low"] + 1) and np.array_equal(ess, f["ess"])
        n = st.num_iters
        assert trk[4, 0] == 0 and (trk[4, 1:n] > 0).all() and (trk[4, n:] == 0).all()
        if sched == 0:
            assert st.schedule in (2, 3) and st.host_reads <= 3
        outs.setdefault(sched, []).append((trk[:4].copy(), n, st.schedule))
    ref, n1, _ = outs[1][0]
    for t, n, _ in outs[0]:
        assert n == n1 and np.array_equal(t, ref)
    # the second call of the same size replays graphs
    assert outs[0][1][2] == 3


def test_low_true_out_of_range_is_rejected():
    b = _gen("vr12")
    r, c = b.coo1()
    lt = np.zeros(b.m, np.int32)
    lt[3] = b.m + 1
    with pytest.raises(openph_amd.OpenPHError) as e:
        openph_amd.ph("pms", r, c, b.m, 8, lt)
    assert e.value.status == -1


def test_autotune_batch_checks_every_candidate_and_returns_a_working_setting():
    say("autotune")
    import torch
    mats = C.batched_mix(32, seed=7)
    ms = np.array([b.m for b in mats], np.int32)
    nn = np.array([b.nnz for b in mats], np.int32)
    rr, cc = zip(*(b.coo1() for b in mats))
    dev = torch.device("cuda:0")
    rows = torch.from_numpy(np.concatenate(rr)).to(dev)
    cols = torch.from_numpy(np.concatenate(cc)).to(dev)
    M = int(ms.sum())
    low = torch.empty(M, dtype=torch.int32, device=dev)
    ess = torch.empty_like(low)
    best, rows_ = openph_amd.autotune_batch_device(ms, nn, rows, cols, low, ess, reps=1)
    assert len(rows_) == 1 + 16 + 4 + 3
    assert all(r["same"] == 1 for r in rows_)
    assert min(r["wall_ms"] for r in rows_) > 0
    st = openph_amd.pms_batch_device(ms, nn, rows, cols, low, ess, options=best)
    torch.cuda.synchronize()
    assert st.lanes == (best.lanes_per_column or st.lanes)
    got = low.cpu().numpy()
    o = 0
    for b in mats:
        ref = oracle.pms(b)
        assert np.array_equal(got[o:o + b.m] - 1, ref.low)
        o += b.m


def test_repeated_size_replays_the_whole_reduction_on_new_contents():
    """A reduction of a size seen before replays one cached graph from the
    set-up to the outputs, keyed on the buffers, not their contents: matrices
    of the same size (VR 2-skeleta of 40 points are all m = 10,700, nnz =
    31,200) written into the same device buffers must each reduce to their
    own oracle result, a malformed one of that size must be refused, and the
    next valid one must again be exact."""
    import torch
    dev = torch.device("cuda:0")
    say("whole-reduction replays")
    mats = [C.vr40(seed) for seed in (0, 1, 2)]
    refs = [oracle.pms(b, snapshot=True) for b in mats]
    m, nnz = mats[0].m, mats[0].nnz
    assert all(b.m == m and b.nnz == nnz for b in mats)
    R = torch.empty(nnz, dtype=torch.int32, device=dev)
    Cc = torch.empty(nnz, dtype=torch.int32, device=dev)
    low = torch.empty(m, dtype=torch.int32, device=dev)
    ess = torch.empty(m, dtype=torch.int32, device=dev)
    bad_r, bad_c = mats[1].coo1()
    bad_r = bad_r.copy()
    bad_r[100], bad_r[101] = bad_r[101], bad_r[100]  # rows out of order in a column
    order = [0, 1, 0, 2, "bad", 1, 2, 0]
    for k, which in enumerate(order):
        if which == "bad":
            R.copy_(torch.from_numpy(bad_r))
            Cc.copy_(torch.from_numpy(bad_c))
            with pytest.raises(openph_amd.OpenPHError) as e:
                openph_amd.pms_device(R, Cc, m, low, ess)
            assert e.value.status == -2, e.value.status  # OPENPH_EUNSORTED
            continue
        r, c = mats[which].coo1()
        R.copy_(torch.from_numpy(r))
        Cc.copy_(torch.from_numpy(c))
        low.fill_(-7)
        st = openph_amd.pms_device(R, Cc, m, low, ess)
        torch.cuda.synchronize()
        ref = refs[which]
        assert np.array_equal(low.cpu().numpy() - 1, ref.low), (k, which)
        assert np.array_equal(ess.cpu().numpy(), ref.ess), (k, which)
        assert st.additions == ref.additions and st.merge_elems == ref.merge_elems, (k, which)
        if k >= 2:  # the size repeats: the automatic schedule replays graphs
            assert st.schedule == O.GRAPH
