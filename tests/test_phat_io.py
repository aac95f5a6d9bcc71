"""CPU tests of the PHAT ascii I/O (numerics/utils/dat2cmo.m:2-48,
cmo2dat.m:2-38, phat.m:46-79) and of the persistence diagram export."""
import math
import os
import sys

import numpy as np
import pytest

from openph_amd import complexes as C
from openph_amd import phat

sys.path.insert(0, os.path.dirname(__file__))
from test_oracle import GOLD, _gen  # noqa: E402


def test_handwritten_file_parses(tmp_path):
    p = tmp_path / "tri.dat"
    p.write_text("# a filled triangle\n0\n0\n\n0\n1 0 1\n1 1 2\n# comment\n1 0 2\n2 5 3 4\n")
    r, c, m, dims = phat.read_boundary(p)
    assert m == 7
    assert dims.tolist() == [0, 0, 0, 1, 1, 1, 2]
    assert r.tolist() == [1, 2, 2, 3, 1, 3, 4, 5, 6]  # faces sorted inside each column
    assert c.tolist() == [4, 4, 5, 5, 6, 6, 7, 7, 7]


@pytest.mark.parametrize("name", ["vr12", "random2", "cubical8"])
def test_boundary_round_trip(name, tmp_path):
    b = _gen(name)
    r, c = b.coo1()
    p = tmp_path / "b.dat"
    phat.write_boundary(p, r, c, b.m, dims=b.dim)
    r2, c2, m2, d2 = phat.read_boundary(p)
    assert m2 == b.m and np.array_equal(r2, r) and np.array_equal(c2, c)
    assert np.array_equal(d2, b.dim)
    # default dimensions: faces - 1 for simplices, 0 for vertices
    phat.write_boundary(p, r, c, b.m)
    if name != "cubical8":
        assert np.array_equal(phat.read_boundary(p)[3], b.dim)


def test_pairs_round_trip(tmp_path):
    f = np.load(os.path.join(GOLD, "vr40.npz"))
    low1 = (f["low"] + 1).astype(np.int32)
    p = tmp_path / "pairs.txt"
    phat.write_pairs(p, low1)
    lines = p.read_text().splitlines()
    assert int(lines[0]) == int((low1 > 0).sum()) == len(lines) - 1
    assert np.array_equal(phat.read_pairs(p, len(low1)), low1)


@pytest.mark.parametrize("n", [10, 15])
def test_diagram_counts_follow_closed_forms(n):
    import oracle
    b = C.vr_2skeleton(C.points_cube(np.random.default_rng(n), n))
    res = oracle.pms(b)
    dg = phat.diagram(res.low + 1, res.ess, b.dim, b.value)
    fin = {d: int(np.isfinite(a[:, 1]).sum()) for d, a in dg.items()}
    inf = {d: int(np.isinf(a[:, 1]).sum()) for d, a in dg.items()}
    assert fin[0] == n - 1 and inf[0] == 1
    assert fin[1] == math.comb(n, 2) - n + 1 and inf[1] == 0
    assert inf[2] == math.comb(n - 1, 3)
    for a in dg.values():
        assert (a[:, 0] <= a[:, 1]).all()


def test_malformed_files_are_rejected(tmp_path):
    p = tmp_path / "bad.dat"
    p.write_text("0\n0\n1 0 0\n")  # a repeated face
    with pytest.raises(ValueError):
        phat.read_boundary(p)
    p.write_text("0\n1 0 1\n")  # a face that is not an earlier column
    with pytest.raises(ValueError):
        phat.read_boundary(p)
    p.write_text("1\n1 4\n")
    with pytest.raises(ValueError):
        phat.read_pairs(p, 3)
