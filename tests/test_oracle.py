"""CPU tests of the oracle and the generators (no GPU).

The oracle is pinned by checks that do not restate its formulas:
dense Z/2 brute force, union-find dimension-0 pairs, closed-form pair counts,
structural invariants of a reduced boundary matrix, the reference's own
compiled merge (oracle/_ref, build container only) and the committed
fixtures. Planted mistakes show each check has teeth.
"""
import json
import math
import os

import numpy as np
import pytest

import oracle
from openph_amd import complexes as C

GOLD = os.path.join(os.path.dirname(__file__), "golden")
INDEX = json.load(open(os.path.join(GOLD, "index.json")))


def _gen(name):
    if name == "vr12":
        return C.vr_2skeleton(C.points_cube(np.random.default_rng(0), 12))
    if name.startswith("random"):
        seed = int(name[len("random"):])
        return C.random_simplicial(np.random.default_rng(100 + seed), 7, 0.7, 0.6, 0.5)
    if name == "vr20_sphere":
        return C.vr_2skeleton(C.points_sphere(np.random.default_rng(3), 20))
    if name == "vr40":
        return C.vr40()
    if name == "torus100":
        return C.torus100()
    if name == "cubical8":
        return C.cubical64(seed=0, side=8)
    if name == "cubical16":
        return C.cubical64(seed=0, side=16)
    raise KeyError(name)


def _sha(b):
    import hashlib
    r, c = b.coo1()
    h = hashlib.sha256()
    h.update(np.int64(b.m).tobytes())
    h.update(r.tobytes())
    h.update(c.tobytes())
    return h.hexdigest()


def invariants(b, low, ess):
    """Facts every reduced boundary matrix satisfies."""
    m = b.m
    pos = low >= 0
    # lows are distinct
    assert np.unique(low[pos]).size == pos.sum()
    # a low row is older than its column and one dimension lower
    j = np.flatnonzero(pos)
    assert (low[j] < j).all()
    assert (b.dim[low[j]] == b.dim[j] - 1).all()
    # a low row is itself a zero column (positive cells reduce to zero)
    assert (low[low[j]] == -1).all()
    # ess = in no pair
    paired = np.zeros(m, bool)
    paired[j] = True
    paired[low[j]] = True
    assert np.array_equal(ess.astype(bool), ~paired)


@pytest.mark.parametrize("name", sorted(INDEX))
def test_fixture_generator_is_stable(name):
    b = _gen(name)
    assert b.m == INDEX[name]["m"] and b.nnz == INDEX[name]["nnz"]
    assert _sha(b) == INDEX[name]["coo_sha256"]


@pytest.mark.parametrize("name", sorted(INDEX))
def test_oracle_matches_fixture(name, oracle_lib):
    f = np.load(os.path.join(GOLD, name + ".npz"))
    b = _gen(name)
    if "rows1" in f:
        r, c = b.coo1()
        assert np.array_equal(r, f["rows1"]) and np.array_equal(c, f["cols1"])
    for algo in (oracle.pms, oracle.standard, oracle.twist):
        res = algo(b)
        assert np.array_equal(res.low, f["low"]), algo.__name__
        assert np.array_equal(res.ess, f["ess"]), algo.__name__
    snap = oracle.pms(b, snapshot=True)
    assert np.array_equal(snap.low, f["low"])
    assert snap.num_iters == INDEX[name]["pms_snapshot"]["num_iters"]
    assert snap.additions == INDEX[name]["pms_snapshot"]["additions"]
    assert np.array_equal(oracle.ph_row(b).low, f["low"])
    invariants(b, f["low"], f["ess"])


@pytest.mark.parametrize("seed", range(12))
def test_oracle_vs_dense_random_complexes(seed, oracle_lib):
    rng = np.random.default_rng(1000 + seed)
    b = C.random_simplicial(rng, 6 + seed % 3, 0.8, 0.7, 0.6)
    dl, de = oracle.dense_reduce(b)
    for algo in (oracle.pms, oracle.standard, oracle.twist):
        res = algo(b)
        assert np.array_equal(res.low, dl), algo.__name__
        assert np.array_equal(res.ess, de), algo.__name__
    assert np.array_equal(oracle.pms(b, snapshot=True).low, dl)
    assert np.array_equal(oracle.ph_row(b).low, dl)
    invariants(b, dl, de)


@pytest.mark.parametrize("n", [8, 15, 25])
def test_vr_closed_forms(n, oracle_lib):
    rng = np.random.default_rng(n)
    b = C.vr_2skeleton(C.points_cube(rng, n))
    assert b.m == n + math.comb(n, 2) + math.comb(n, 3)
    assert b.nnz == 2 * math.comb(n, 2) + 3 * math.comb(n, 3)
    res = oracle.pms(b)
    low, ess = res.low, res.ess
    pos = np.flatnonzero(low >= 0)
    d = b.dim[low[pos]]
    assert (d == 0).sum() == n - 1                       # vertex-edge pairs
    assert (d == 1).sum() == math.comb(n, 2) - n + 1     # edge-triangle pairs
    assert ess.sum() == 1 + math.comb(n - 1, 3)          # beta0 = 1, beta2
    # dimension-0 pairs by union-find and the elder rule
    uf = oracle.dim0_pairs_vr(b)
    got = {int(j): int(low[j]) for j in pos if b.dim[j] == 1}
    assert got == uf


def test_torus_dim0_and_counts(oracle_lib):
    b = C.torus100()
    f = np.load(os.path.join(GOLD, "torus100.npz"))
    low = f["low"]
    pos = np.flatnonzero(low >= 0)
    assert {int(j): int(low[j]) for j in pos if b.dim[j] == 1} == oracle.dim0_pairs_vr(b)
    assert f["ess"].sum() == 1 + math.comb(99, 3)


@pytest.mark.parametrize("side", [4, 8])
def test_cubical_closed_forms(side, oracle_lib):
    b = C.cubical64(seed=side, side=side)
    n = 2 * side - 1
    assert b.m == n ** 3
    e = 3 * (side - 1) * side ** 2
    f = 3 * (side - 1) ** 2 * side
    c = (side - 1) ** 3
    assert b.nnz == 2 * e + 4 * f + 6 * c
    res = oracle.pms(b)
    # a full grid is contractible: one essential class, the oldest vertex
    ess = np.flatnonzero(res.ess == 1)
    assert ess.tolist() == [0]
    assert b.dim[0] == 0
    invariants(b, res.low, res.ess)


def test_planted_mistakes_are_caught(oracle_lib):
    """Plausible mistakes in a result must fail the checks used above."""
    b = C.vr_2skeleton(C.points_cube(np.random.default_rng(7), 10))
    dl, de = oracle.dense_reduce(b)
    good = oracle.pms(b)
    assert np.array_equal(good.low, dl)
    # (1) 1-based output mistaken for 0-based
    assert not np.array_equal(good.low + 1, dl)
    # (2) no reduction at all (initial lows)
    init = np.array([b.rows[b.colptr[j + 1] - 1] if b.colptr[j + 1] > b.colptr[j] else -1
                     for j in range(b.m)], np.int32)
    assert not np.array_equal(init, dl)
    with pytest.raises(AssertionError):
        invariants(b, init, de)
    # (3) clearing applied to the wrong index (row instead of column)
    bad = dl.copy()
    j = np.flatnonzero(dl >= 0)[-1]
    bad[dl[j]] = dl[j]
    with pytest.raises(AssertionError):
        invariants(b, bad, de)
    # (4) ess that ignores the births of pairs
    bad_ess = de.copy()
    bad_ess[dl[j]] = 1
    with pytest.raises(AssertionError):
        invariants(b, dl, bad_ess)
    # (5) a merge without cancellation (union instead of xor): the reduced
    # lows can only stay the same or grow, which the dense check rejects
    M = np.zeros((b.m, b.m), np.uint8)
    for c in range(b.m):
        M[b.rows[b.colptr[c]:b.colptr[c + 1]], c] = 1
    lows_union = []
    owner = {}
    for c in range(b.m):
        col = M[:, c].copy()
        for _ in range(b.m):
            nz = np.flatnonzero(col)
            if nz.size == 0 or int(nz[-1]) not in owner:
                break
            col |= M[:, owner[int(nz[-1])]]
            col[nz[-1]] = 0  # drops only the low, keeps everything else
        nz = np.flatnonzero(col)
        lows_union.append(int(nz[-1]) if nz.size else -1)
        if nz.size:
            owner.setdefault(int(nz[-1]), c)
    assert not np.array_equal(np.array(lows_union, np.int32), dl)


@pytest.mark.skipif(not os.path.isdir("/root/reference/src/cuda"),
                    reason="reference sources are only in the build container")
@pytest.mark.parametrize("name", ["vr12", "random0", "vr20_sphere", "vr40", "cubical8"])
def test_oracle_vs_reference_build(name, oracle_lib):
    oracle.build_ref()
    b = _gen(name)
    res = oracle.standard(b)
    mx = max(1, int(np.diff(b.colptr).max()))
    cw = -(-max(res.max_len, mx) // mx) + 1
    for alg in ("standard", "twist"):
        rl, re = oracle.ref_run(alg, b, cw)
        assert np.array_equal(rl, res.low)
        assert np.array_equal(re, res.ess)


def test_oracle_slot_overflow_is_an_error(oracle_lib):
    """The fixed-slot restatement reports overflow (the reference exits,
    src/cuda/algorithms/kernels_host.cu:21-24)."""
    b = C.vr40()
    lib = oracle._lib()
    colptr, rows = oracle._csc(b)
    low = np.empty(b.m, np.int32)
    ess = np.empty(b.m, np.int32)
    st = np.zeros(4, np.int64)
    rc = lib.orc_twist(b.m, colptr, rows, 1, low, ess, st)
    assert rc == oracle.ORC_EOVERFLOW


@pytest.mark.parametrize("name", ["vr12", "random0", "random2", "vr20_sphere", "vr40"])
def test_oracle_trackers(name, oracle_lib):
    """The integers behind the reference's trackers (tracker.cu:92-134) per
    iteration, pinned at both ends by closed forms: before Phase 0 every
    column is unreduced and essential and the low distance is that of the
    input's last rows; after the last iteration low equals low_true, nothing
    is unreduced (update_classes_final classifies every column) and the
    essential estimate is exact. In between, classes and ess only ever move
    away from 0 and 1, so both counts never increase."""
    f = np.load(os.path.join(GOLD, name + ".npz"))
    b = _gen(name)
    lt = f["low"].astype(np.int32)
    res = oracle.pms(b, snapshot=True, low_true=lt)
    trk = res.trk
    assert trk.shape == (res.num_iters, 4)
    m = b.m
    counts = np.diff(b.colptr)
    low0 = np.where(counts > 0, b.rows[np.maximum(b.colptr[1:] - 1, 0)], -1)
    d0 = np.abs(low0.astype(np.int64) - lt)
    assert trk[0].tolist() == [d0.sum(), d0.max(), m, m]
    paired = np.zeros(m, bool)
    j = np.flatnonzero(lt >= 0)
    paired[j] = True
    paired[lt[j]] = True
    assert trk[-1].tolist() == [0, 0, 0, int((~paired).sum())]
    assert (np.diff(trk[:, 2]) <= 0).all() and (np.diff(trk[:, 3]) <= 0).all()
    # a wrong low_true shows up in the distances
    lt2 = lt.copy()
    lt2[j[0]] = -1
    bad = oracle.pms(b, snapshot=True, low_true=lt2).trk
    assert bad[-1][0] == lt[j[0]] + 1 and bad[-1][1] == lt[j[0]] + 1


def _row_terminated(b, low):
    """ph_row's end state: for every row, no column of its leftmost column's
    pseudo-dimension (initial length) shares that low."""
    plen = np.diff(b.colptr)
    first = {}
    for j in np.flatnonzero(low >= 0):
        first.setdefault(int(low[j]), int(j))
    for j in np.flatnonzero(low >= 0):
        p = first[int(low[j])]
        if j != p and plen[j] == plen[p]:
            return False
    return True


@pytest.mark.parametrize("seed", range(6))
def test_row_algorithms_on_random_matrices(seed, oracle_lib):
    """ph_row_parallel (ph_row_parallel.cu:20-84) reduces every column, so
    its low is the dense reduction's, and its ess (ess_hat with no arglow) is
    0 exactly at the lows. ph_row (ph_row.cu:38-121) only reduces inside the
    pivot's pseudo-dimension: on a matrix whose equal-low columns differ in
    length it stops short of the reduction, in the state _row_terminated
    describes, with lows never above the initial ones and ess <= the exact
    estimate."""
    b = C.random_upper(np.random.default_rng(2000 + seed), 200)
    dl, de = oracle.dense_reduce(b)
    par = oracle.ph_row(b, parallel=True)
    assert np.array_equal(par.low, dl)
    at_low = np.zeros(b.m, bool)
    at_low[dl[dl >= 0]] = True
    assert np.array_equal(par.ess, (~at_low).astype(np.int32))
    row = oracle.ph_row(b)
    assert not np.array_equal(row.low, dl)  # this matrix tells the two apart
    assert _row_terminated(b, row.low)
    counts = np.diff(b.colptr)
    low0 = np.where(counts > 0, b.rows[np.maximum(b.colptr[1:] - 1, 0)], -1)
    assert (row.low <= low0).all()
    assert not _row_terminated(b, low0)
    # the same additions on a complex, where equal lows mean equal lengths
    c = C.random_simplicial(np.random.default_rng(seed), 7, 0.8, 0.7, 0.6)
    cl, ce = oracle.dense_reduce(c)
    r = oracle.ph_row(c)
    assert np.array_equal(r.low, cl) and (r.ess <= ce).all()
    assert r.additions == oracle.ph_row(c, parallel=True).additions


@pytest.mark.parametrize("name", sorted(INDEX))
def test_twist_on_growable_columns_equals_fixed_slot_twist(name, oracle_lib):
    """orc_twist_dyn (the CPU baseline for matrices whose fixed-slot layout
    does not fit in memory) takes the same columns in the same order and adds
    the same pivots as orc_twist: outputs and counters equal."""
    f = np.load(os.path.join(GOLD, name + ".npz"))
    b = _gen(name)
    d = oracle.twist_dyn(b)
    t = oracle.twist(b)
    assert np.array_equal(d.low, f["low"]) and np.array_equal(d.ess, f["ess"])
    assert (d.additions, d.merge_elems, d.max_len, d.iters) == \
        (t.additions, t.merge_elems, t.max_len, t.iters)


def test_digest_fixtures_are_consistent():
    """The digest fixtures of the batched and served workloads name the
    matrices the generators still produce (checked on the small members)."""
    import hashlib
    import json

    def coo_digest(b):
        r, c = b.coo1()
        h = hashlib.sha256()
        h.update(np.int64(b.m).tobytes())
        h.update(np.ascontiguousarray(r, dtype="<i4").tobytes())
        h.update(np.ascontiguousarray(c, dtype="<i4").tobytes())
        return h.hexdigest()[:16]

    mix = json.load(open(os.path.join(GOLD, "batched512.json")))["members"]
    assert len(mix) == 512 and sum(x["m"] for x in mix) == 14757723
    mats = C.batched_mix(8, seed=0)  # the first children of the same seed sequence
    for b, g in zip(mats, mix):
        assert b.m == g["m"] and coo_digest(b) == g["coo"]
        r = oracle.pms(b, snapshot=True)
        assert hashlib.sha256(r.low.astype("<i4").tobytes()).hexdigest()[:16] == g["low"]
    served = json.load(open(os.path.join(GOLD, "served17.json")))["requests"]
    kinds = [(s["kind"], s["size"]) for s in served]
    assert ("points", 190) in kinds and ("field", 85) in kinds and ("field", 81) in kinds
