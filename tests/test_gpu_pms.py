"""GPU parity tests: the HIP pms engine through the C ABI against the oracle.

Bar: bit-exact low and ess. Where the oracle runs the same schedule (its
snapshot mode publishes Phase II pivots after the pass, as the GPU does) the
iteration count, the number of column additions and the merged elements
must match too.
"""
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
from test_oracle import _gen, invariants  # noqa: E402


def say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    openph_amd.lib()


@pytest.mark.parametrize("name", sorted(INDEX))
def test_pms_matches_golden(name):
    say("case", name)
    f = np.load(os.path.join(GOLD, name + ".npz"))
    b = _gen(name)
    low, ess, st = openph_amd.reduce(b, "pms")
    assert np.array_equal(low, f["low"])
    assert np.array_equal(ess, f["ess"])
    snap = INDEX[name]["pms_snapshot"]
    # a pool overflow cuts a Phase II pass short and changes the schedule
    assert st.overflows == 0
    assert st.num_iters == snap["num_iters"]
    assert st.additions == snap["additions"]
    assert st.merge_elems == snap["merge_elems"]
    assert st.m == b.m


SHAPES = [(1, 0), (4, 32), (4, 64), (8, 32), (8, 64), (8, 72), (8, 80), (8, 96), (8, 128), (16, 64),
          (16, 128), (16, 256), (32, 256), (32, 512), (64, 512), (64, 1024)]


@pytest.mark.parametrize("lanes,wlds", SHAPES)
@pytest.mark.parametrize("name", ["torus100", "cubical16", "vr20_sphere"])
def test_every_lane_configuration_is_exact(name, lanes, wlds):
    """The Phase II schedule does not depend on the lane configuration, so
    the counters must match the oracle's snapshot schedule as well."""
    say("lanes", name, lanes, wlds)
    f = np.load(os.path.join(GOLD, name + ".npz"))
    opt = openph_amd.Options()
    opt.lanes_per_column = lanes
    opt.group_lds = wlds
    low, ess, st = openph_amd.reduce(_gen(name), "pms", opt)
    assert st.lanes == lanes
    assert np.array_equal(low, f["low"])
    assert np.array_equal(ess, f["ess"])
    snap = INDEX[name]["pms_snapshot"]
    assert st.overflows == 0
    assert st.additions == snap["additions"]
    assert st.merge_elems == snap["merge_elems"]
    assert st.num_iters == snap["num_iters"]


@pytest.mark.parametrize("lanes", [4, 8])
def test_lds_overflow_path_is_exact(lanes):
    """Cubical columns outgrow a lane group's LDS buffers; they are handed to
    the block kernel, which must continue the same chain."""
    say("lds overflow", lanes)
    b = C.cubical64(seed=40, side=40)
    ref = oracle.pms(b, snapshot=True)
    opt = openph_amd.Options()
    opt.lanes_per_column = lanes
    low, ess, st = openph_amd.reduce(b, "pms", opt)
    assert st.handoffs > 0
    assert np.array_equal(low, ref.low)
    assert np.array_equal(ess, ref.ess)
    assert st.additions == ref.additions


@pytest.mark.parametrize("algo", openph_amd.ALGORITHMS)
def test_every_algstr_gives_the_reference_low_and_ess(algo):
    say("algo", algo)
    f = np.load(os.path.join(GOLD, "vr40.npz"))
    b = C.vr40()
    r, c = b.coo1()
    out = openph_amd.ph(algo, r, c, b.m, 8, np.zeros(b.m, np.int32))
    assert len(out) == 8
    low1, ess = out[0], out[1]
    assert np.array_equal(low1, f["low"] + 1)
    if algo.startswith("ph_row"):
        # the row algorithms' own estimates (ph_row.cu:14-21, ess_hat without
        # arglow in ph_row_parallel.cu:69)
        ess_ref = oracle.ph_row(b, parallel=algo == "ph_row_parallel").ess
        assert not np.array_equal(ess_ref, f["ess"])
    else:
        ess_ref = f["ess"]
    assert np.array_equal(ess, ess_ref)
    assert out[7][0] >= 1


@pytest.mark.parametrize("side", [24, 40])
def test_cubical_vs_oracle(side):
    say("cubical", side)
    b = C.cubical64(seed=side, side=side)
    ref = oracle.pms(b, snapshot=True)
    low, ess, st = openph_amd.reduce(b, "pms")
    assert np.array_equal(low, ref.low)
    assert np.array_equal(ess, ref.ess)
    assert st.additions == ref.additions
    assert st.num_iters == ref.num_iters


@pytest.mark.slow
def test_cubical64_full_size():
    say("cubical64")
    b = C.cubical64()
    ref = oracle.pms(b, snapshot=True)
    low, ess, st = openph_amd.reduce(b, "pms")
    assert np.array_equal(low, ref.low)
    assert np.array_equal(ess, ref.ess)
    # contractible grid: a single essential class
    assert int(ess.sum()) == 1
    # 8 lanes with 80-entry buffers: the handed-off columns go to one wave
    # each with 1024-entry buffers, and those that outgrow them merge in the
    # pool
    opt = openph_amd.Options()
    opt.lanes_per_column = 8
    low, ess, st = openph_amd.reduce(b, "pms", opt)
    assert st.handoffs > 0 and st.spills > 0
    assert np.array_equal(low, ref.low)
    assert np.array_equal(ess, ref.ess)
    assert st.additions == ref.additions


def test_device_entry_point_with_torch_buffers():
    import torch
    say("device entry")
    b = C.torus100()
    f = np.load(os.path.join(GOLD, "torus100.npz"))
    r, c = b.coo1()
    dr = torch.from_numpy(r).cuda()
    dc = torch.from_numpy(c).cuda()
    low = torch.full((b.m,), -7, dtype=torch.int32, device="cuda")
    ess = torch.full((b.m,), -7, dtype=torch.int32, device="cuda")
    for _ in range(3):  # workspace reuse across calls
        st = openph_amd.pms_device(dr, dc, b.m, low, ess)
        torch.cuda.synchronize()
        assert np.array_equal(low.cpu().numpy(), f["low"] + 1)
        assert np.array_equal(ess.cpu().numpy(), f["ess"])
        assert st.num_iters == INDEX["torus100"]["pms_snapshot"]["num_iters"]


def test_small_pool_compacts_and_stays_exact():
    say("small pool")
    b = C.torus100()
    f = np.load(os.path.join(GOLD, "torus100.npz"))
    opt = openph_amd.Options()
    opt.slot_elems = 1
    opt.pool_elems = b.nnz + b.m + 4096
    low, ess, st = openph_amd.reduce(b, "pms", opt)
    assert st.compactions > 0 and st.overflows > 0
    assert np.array_equal(low, f["low"])
    assert np.array_equal(ess, f["ess"])
    # suspended passes run again before their iteration ends
    snap = INDEX["torus100"]["pms_snapshot"]
    assert st.additions == snap["additions"]
    assert st.num_iters == snap["num_iters"]


def test_small_pool_with_handoffs_compacts_and_stays_exact():
    """Compactions of a small pool while handed-off columns are merged by
    the wave teams of the long-column pass."""
    say("small pool, handoffs")
    b = C.cubical64(seed=40, side=40)
    ref = oracle.pms(b, snapshot=True)
    opt = openph_amd.Options()
    opt.lanes_per_column = 8
    opt.slot_elems = 1
    opt.pool_elems = b.nnz + b.m + 4096
    low, ess, st = openph_amd.reduce(b, "pms", opt)
    assert st.handoffs > 0
    assert st.compactions > 0
    assert np.array_equal(low, ref.low)
    assert np.array_equal(ess, ref.ess)
    assert st.additions == ref.additions
    assert st.num_iters == ref.num_iters


_SUSPEND = {}


def _suspend_case():
    if not _SUSPEND:
        b = C.cubical_lower_star(C.smoothed_field(np.random.default_rng(11), 43))
        _SUSPEND["b"] = b
        _SUSPEND["ref"] = oracle.pms(b, snapshot=True)
    return _SUSPEND["b"], _SUSPEND["ref"]


@pytest.mark.parametrize("schedule", [1, 2, 3])
@pytest.mark.parametrize("lanes,wlds", [(0, 0), (8, 32), (32, 256), (1, 0)])
def test_phase2_out_of_pool_is_resumed_with_the_oracles_counters(schedule, lanes, wlds):
    """A Phase II pass that runs out of pool is suspended, the pool compacted
    and the same pass run again before the iteration ends, so the additions,
    the merged elements and the iteration count stay the oracle's (pms.cu's
    loop never sees a full pool)."""
    say("suspended pass", schedule, lanes, wlds)
    b, ref = _suspend_case()
    opt = openph_amd.Options()
    opt.schedule = schedule
    opt.lanes_per_column = lanes
    opt.group_lds = wlds
    opt.pool_elems = b.nnz + b.m + 2048
    low, ess, st = openph_amd.reduce(b, "pms", opt)
    assert st.overflows > 0 and st.compactions > 0
    assert np.array_equal(low, ref.low)
    assert np.array_equal(ess, ref.ess)
    assert st.additions == ref.additions
    assert st.merge_elems == ref.merge_elems
    assert st.num_iters == ref.num_iters


@pytest.mark.parametrize("seed", range(6))
def test_tiny_matrix_with_a_pool_below_its_layout(seed):
    """A pool asked below the initial layout starts full; its compactions
    grow it to leave room, whatever the matrix's size."""
    say("tiny pool", seed)
    rng = np.random.default_rng(100 + seed)
    b = C.random_simplicial(rng, 11, p_edge=0.6)
    ref = oracle.pms(b, snapshot=True)
    opt = openph_amd.Options()
    opt.pool_elems = b.nnz + b.m + 1
    low, ess, st = openph_amd.reduce(b, "pms", opt)
    assert np.array_equal(low, ref.low)
    assert np.array_equal(ess, ref.ess)
    assert st.additions == ref.additions
    assert st.num_iters == ref.num_iters


@pytest.mark.parametrize("extra", [0, 1024, 65536])
def test_pool_budget_too_small_is_an_error_not_a_wrong_answer(extra):
    """With the pool capped near the input size the reduction either finishes
    exactly (compacting as it goes) or reports ENOMEM; it never returns a
    wrong answer."""
    say("pool budget", extra)
    b = C.torus100()
    f = np.load(os.path.join(GOLD, "torus100.npz"))
    opt = openph_amd.Options()
    opt.pool_elems = b.nnz + extra
    opt.pool_max_elems = b.nnz + extra
    try:
        low, ess, st = openph_amd.reduce(b, "pms", opt)
    except openph_amd.OpenPHError as e:
        assert e.status == -4
        return
    assert np.array_equal(low, f["low"])
    assert np.array_equal(ess, f["ess"])


def test_sizes_in_a_row_reuse_the_workspace():
    say("sizes in a row")
    for name in ["vr40", "vr12", "torus100", "random1", "cubical8"]:
        f = np.load(os.path.join(GOLD, name + ".npz"))
        low, ess, _ = openph_amd.reduce(_gen(name), "pms")
        assert np.array_equal(low, f["low"]), name
        assert np.array_equal(ess, f["ess"]), name


def test_edge_cases_empty_and_vertices_only():
    say("edge cases")
    b = C.Boundary(colptr=np.zeros(1, np.int64), rows=np.zeros(0, np.int32),
                   dim=np.zeros(0, np.int32), value=np.zeros(0))
    low, ess, st = openph_amd.reduce(b, "pms")
    assert low.size == 0 and ess.size == 0
    b = C.Boundary(colptr=np.zeros(6, np.int64), rows=np.zeros(0, np.int32),
                   dim=np.zeros(5, np.int32), value=np.zeros(5))
    low, ess, st = openph_amd.reduce(b, "pms")
    assert (low == -1).all() and (ess == 1).all()


@pytest.mark.parametrize("kind", ["row_ge_col", "unsorted_cols", "unsorted_rows", "out_of_range"])
def test_malformed_input_is_rejected(kind):
    say("malformed", kind)
    # a triangle: vertices 1..3, edges 4..6, face 7
    r = np.array([1, 2, 1, 3, 2, 3, 4, 5, 6], np.int32)
    c = np.array([4, 4, 5, 5, 6, 6, 7, 7, 7], np.int32)
    if kind == "row_ge_col":
        r = r.copy(); r[0] = 4
    elif kind == "unsorted_cols":
        c = c.copy(); c[2], c[6] = c[6], c[2]
    elif kind == "unsorted_rows":
        r = r.copy(); r[0], r[1] = r[1], r[0]
    else:
        r = r.copy(); r[-1] = 99
    with pytest.raises(openph_amd.OpenPHError) as e:
        openph_amd.ph("pms", r, c, 7, 1, None)
    assert e.value.status in (-1, -2)
    # the same call with the valid matrix still works afterwards
    r0 = np.array([1, 2, 1, 3, 2, 3, 4, 5, 6], np.int32)
    c0 = np.array([4, 4, 5, 5, 6, 6, 7, 7, 7], np.int32)
    out = openph_amd.ph("pms", r0, c0, 7, 1, None)
    assert out[0].tolist() == [0, 0, 0, 2, 3, 0, 6]
    assert out[1].tolist() == [1, 0, 0, 0, 0, 0, 0]


@pytest.mark.parametrize("kind", ["row_ge_col", "unsorted_rows", "out_of_range"])
def test_same_size_malformed_input_is_rejected_without_a_setup_read(kind):
    """A reduction of the size the workspace reduced last does not wait for
    its set-up (the Phase II shape is the last one's); a malformed matrix of
    that size must still be refused with its code, through a guard kernel
    that makes it inert, with the outputs left as given, and the valid
    matrix reduced next must be exact."""
    import torch
    say("same-size malformed", kind)
    b = _gen("vr40")
    f = np.load(os.path.join(GOLD, "vr40.npz"))
    r, c = b.coo1()
    low0, ess0, _ = openph_amd.reduce(b, "pms")  # the workspace's last size
    assert np.array_equal(low0, f["low"])
    rr = r.copy()
    if kind == "row_ge_col":
        rr[7] = c[7]
        want = -1
    elif kind == "unsorted_rows":
        j = int(np.flatnonzero(c[1:] == c[:-1])[10])
        rr[j], rr[j + 1] = rr[j + 1], rr[j]
        want = -2
    else:
        rr[-1] = b.m + 5
        want = -1
    with pytest.raises(openph_amd.OpenPHError) as e:
        openph_amd.ph("pms", rr, c, b.m, 1, None)
    assert e.value.status == want
    # the device entry, twice at this size: outputs keep the caller's values
    dev = torch.device("cuda:0")
    dr, dc = torch.from_numpy(rr).to(dev), torch.from_numpy(c).to(dev)
    for _ in range(2):
        low = torch.full((b.m,), -7, dtype=torch.int32, device=dev)
        ess = torch.full((b.m,), -7, dtype=torch.int32, device=dev)
        with pytest.raises(openph_amd.OpenPHError) as e:
            openph_amd.pms_device(dr, dc, b.m, low, ess)
        assert e.value.status == want
        torch.cuda.synchronize()
        assert int(low.min()) == -7 == int(low.max()) and int(ess.min()) == -7
    low1, ess1, st = openph_amd.reduce(b, "pms")
    assert np.array_equal(low1, f["low"]) and np.array_equal(ess1, f["ess"])
    assert st.host_reads == 1


def test_dropin_openph_matches_openph_m_fields():
    say("openph()")
    b = C.vr40()
    r, c = b.coo1()
    out, t = openph_amd.openph(r, c, b.m, "pms", 8, np.zeros(b.m, np.int32))
    f = np.load(os.path.join(GOLD, "vr40.npz"))
    assert np.array_equal(out["low"], f["low"] + 1)
    for k in ("ess", "err_linf", "err_lone", "err_redu", "err_ess", "time_track",
              "num_iters", "time_inner", "m"):
        assert k in out
    assert t > 0


# ---------------------------------------------------------------------------
# batches (block-diagonal offsets)
# ---------------------------------------------------------------------------

def test_batch_of_golden_matrices_matches_each_alone():
    say("batch golden")
    names = ["vr12", "random0", "random1", "random2", "random3", "vr20_sphere", "cubical8",
             "vr40", "torus100"]
    mats = [_gen(n) for n in names]
    out, st = openph_amd.pms_batch(mats)
    for n, (low, ess) in zip(names, out):
        f = np.load(os.path.join(GOLD, n + ".npz"))
        assert np.array_equal(low, f["low"]), n
        assert np.array_equal(ess, f["ess"]), n
    # matrices do not interact: the work is the sum of the parts
    assert st.additions == sum(INDEX[n]["pms_snapshot"]["additions"] for n in names)
    assert st.merge_elems == sum(INDEX[n]["pms_snapshot"]["merge_elems"] for n in names)
    assert st.num_iters == max(INDEX[n]["pms_snapshot"]["num_iters"] for n in names)


def test_batch_mix_subset_vs_oracle():
    say("batch mix")
    mats = C.batched_mix(24, seed=5)
    out, st = openph_amd.pms_batch(mats)
    adds = 0
    for b, (low, ess) in zip(mats, out):
        ref = oracle.pms(b, snapshot=True)
        assert np.array_equal(low, ref.low), b.name
        assert np.array_equal(ess, ref.ess), b.name
        adds += ref.additions
    assert st.additions == adds


def test_batch_device_with_empty_members():
    import torch
    say("batch device")
    empty = C.Boundary(colptr=np.zeros(1, np.int64), rows=np.zeros(0, np.int32),
                       dim=np.zeros(0, np.int32), value=np.zeros(0))
    verts = C.Boundary(colptr=np.zeros(4, np.int64), rows=np.zeros(0, np.int32),
                       dim=np.zeros(3, np.int32), value=np.zeros(3))
    mats = [_gen("vr12"), empty, verts, _gen("random2"), empty]
    rr, cc = zip(*(b.coo1() for b in mats))
    rows = torch.from_numpy(np.concatenate(rr)).cuda()
    cols = torch.from_numpy(np.concatenate(cc)).cuda()
    M = sum(b.m for b in mats)
    low = torch.full((M,), -9, dtype=torch.int32, device="cuda")
    ess = torch.full((M,), -9, dtype=torch.int32, device="cuda")
    openph_amd.pms_batch_device([b.m for b in mats], [b.nnz for b in mats], rows, cols, low, ess)
    torch.cuda.synchronize()
    low, ess = low.cpu().numpy(), ess.cpu().numpy()
    o = 0
    for b in mats:
        ref = oracle.pms(b)
        assert np.array_equal(low[o:o + b.m] - 1, ref.low)
        assert np.array_equal(ess[o:o + b.m], ref.ess)
        o += b.m


def test_batch_rejects_index_outside_its_matrix():
    say("batch invalid")
    a, b = _gen("random0"), _gen("random1")
    r, c = b.coo1()
    c = c.copy()
    c[-1] = b.m + 3  # points into the next matrix's range
    bad = C.Boundary(colptr=b.colptr, rows=b.rows, dim=b.dim, value=b.value)
    ms = np.array([b.m, a.m], np.int32)
    nnzs = np.array([b.nnz, a.nnz], np.int32)
    ra, ca = a.coo1()
    rows = np.concatenate([r, ra]).astype(np.int32)
    cols = np.concatenate([c, ca]).astype(np.int32)
    low = np.zeros(int(ms.sum()), np.int32)
    ess = np.zeros(int(ms.sum()), np.int32)
    import ctypes
    L = openph_amd.lib()
    ip = lambda x: x.ctypes.data_as(ctypes.POINTER(ctypes.c_int))
    rc = L.openph_pms_batch(2, ip(ms), ip(nnzs), ip(rows), ip(cols), None, ip(low), ip(ess), None)
    assert rc == -1
    del bad


def _polygons_boundary(isolated, sizes, faces_per_cycle=3):
    """A 2-dimensional cell complex in filtration order: `isolated` vertices
    with no cofaces, then for each k in `sizes` a cycle of k vertices and k
    edges, all vertices first, then all edges, then `faces_per_cycle` 2-cells
    per cycle, each bounded by all k edges of its cycle (a boundary matrix:
    every vertex lies on two edges of its cycle). Input columns run from 0 to
    max(sizes) entries, after a run of `isolated` + sum(sizes) empty ones."""
    nv = isolated + sum(sizes)
    vbase, v = [], isolated
    for k in sizes:
        vbase.append(v)
        v += k
    cols = [[] for _ in range(nv)]
    ebase, e = [], nv
    for k, vb in zip(sizes, vbase):
        ebase.append(e)
        for i in range(k):
            a, b = vb + i, vb + (i + 1) % k
            cols.append(sorted((a, b)))
            e += 1
    for k, eb in zip(sizes, ebase):
        for _ in range(faces_per_cycle):
            cols.append(list(range(eb, eb + k)))
    counts = np.array([len(c) for c in cols], dtype=np.int64)
    colptr = np.concatenate([[0], np.cumsum(counts)]).astype(np.int64)
    rows = np.array([r for c in cols for r in c], dtype=np.int32)
    dim = np.array([0] * nv + [1] * (e - nv) + [2] * (len(cols) - e), dtype=np.int32)
    return C.Boundary(colptr=colptr, rows=rows, dim=dim,
                      value=np.arange(len(cols), dtype=np.float64), name="polygons")


def test_setup_long_empty_runs_and_long_input_columns():
    """The set-up's column extents and pool layout where a wave's columns
    have no entries at all (thousands of empty columns in a row) and where
    one wave's columns hold many windows of entries (input columns of up to
    2,500 entries): low, ess and the counters equal the oracle's."""
    say("polygons")
    b = _polygons_boundary(9000, [5, 70, 300, 1000, 2500])
    ref = oracle.pms(b, snapshot=True)
    low, ess, st = openph_amd.reduce(b, "pms")
    assert np.array_equal(low, ref.low)
    assert np.array_equal(ess, ref.ess)
    assert st.num_iters == ref.num_iters
    assert st.additions == ref.additions
    assert st.merge_elems == ref.merge_elems
