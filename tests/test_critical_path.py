"""The critical path of Phase II, per iteration, from the oracle's iteration
statistics: the longest chain of dependent additions of one column, under
the snapshot schedule this build runs (pivots found in a Phase II pass are
published after it) and under the reference's own immediate publication
(pms_functions.cu:101, arglow written inside the pass; the oracle runs the
columns in index order, so each sees every pivot published before it).

The sum of those chains bounds a reduction's Phase II from below at any
per-step latency; these values are the table the design's Phase II section
rests on. Immediate publication does not shorten it: on cubical 64^3 it is
longer (4,831 against 4,771 steps) and takes 1.9 times the additions; on
Torus-100 it saves 7 of 221 steps. Both schedules reach the same low and ess."""
import numpy as np

import oracle
from openph_amd import complexes as C


def _cubical64():
    return C.cubical_lower_star(C.smoothed_field(np.random.default_rng(0), 64))


def _chains(b, snapshot):
    r = oracle.pms(b, iter_stats=True, snapshot=snapshot)
    return r, [int(x) for x in r.iter_stats[:, 4]], int(r.iter_stats[:, 0].sum())


def test_torus100_critical_path():
    b = C.torus100()
    rs, snap, add_s = _chains(b, True)
    ri, imm, add_i = _chains(b, False)
    assert snap == [55, 35, 44, 14, 14, 38, 21] and sum(snap) == 221
    assert imm == [55, 28, 44, 14, 14, 38, 21] and sum(imm) == 214
    assert rs.num_iters == ri.num_iters == 8
    assert (add_s, add_i) == (3031373, 3031371)
    assert np.array_equal(rs.low, ri.low) and np.array_equal(rs.ess, ri.ess)


def test_cubical64_critical_path():
    b = _cubical64()
    rs, snap, add_s = _chains(b, True)
    ri, imm, add_i = _chains(b, False)
    assert snap == [253, 330, 349, 622, 368, 241, 565, 254, 556, 302, 643, 288]
    assert sum(snap) == 4771
    assert imm == [413, 273, 622, 344, 368, 565, 203, 254, 556, 302, 643, 288]
    assert sum(imm) == 4831
    assert rs.num_iters == ri.num_iters == 13
    assert (add_s, add_i) == (313523, 594314)
    assert np.array_equal(rs.low, ri.low) and np.array_equal(rs.ess, ri.ess)
