"""Randomised parity on the GPU: VR 2-skeleta, cubical lower-star
filtrations and random simplicial complexes of random sizes, reduced with
random lane shapes, schedules, slot sizes, chunk lengths and initial pools
(small pools compact and suspend Phase II passes), alone and in random
batches. low, ess, the additions, the merged elements and the iteration
count must equal the oracle's snapshot pms (pms.cu:1-118 with the pivots of
each iteration's start)."""
import numpy as np
import pytest

import oracle
import openph_amd
from openph_amd import complexes as C

pytestmark = pytest.mark.gpu

SHAPES = [(4, 32), (4, 64), (8, 32), (8, 64), (8, 72), (8, 80), (8, 96), (8, 128), (16, 64),
          (16, 128), (16, 256), (32, 256), (32, 512), (64, 512), (64, 1024)]


def say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    openph_amd.lib()


def _matrix(rng):
    k = rng.integers(0, 4)
    if k == 0:
        n = int(rng.integers(8, 120))
        f = (C.points_cube, C.points_torus, C.points_sphere, C.points_gmm)[rng.integers(0, 4)]
        p = f(rng, n)
        if rng.random() < 0.3:
            p = np.round(p * 4) / 4  # ties and repeated points
        return f"vr {f.__name__} n={n}", C.vr_2skeleton(p)
    if k == 1:
        side = int(rng.integers(3, 33))
        fld = C.smoothed_field(rng, side, sigma=float(rng.uniform(0.5, 3)))
        if rng.random() < 0.3:
            fld = np.round(fld * 3) / 3
        return f"cubical side={side}", C.cubical_lower_star(fld)
    if k == 2:
        nv = int(rng.integers(5, 40))
        return f"simplicial nv={nv}", C.random_simplicial(rng, nv, p_edge=float(rng.uniform(0.2, 0.8)))
    side = int(rng.integers(36, 49))
    return f"cubical side={side}", C.cubical_lower_star(C.smoothed_field(rng, side))


def _options(rng, b):
    o = openph_amd.Options()
    r = rng.random()
    if 0.3 <= r < 0.4:
        o.lanes_per_column = 1
    elif r >= 0.4:
        o.lanes_per_column, o.group_lds = SHAPES[rng.integers(0, len(SHAPES))]
    o.schedule = int(rng.integers(0, 4))
    if rng.random() < 0.3:
        o.slot_elems = int(rng.choice([1, 8, 32, 64]))
    if rng.random() < 0.25:
        o.pool_elems = int(b.nnz + b.m + rng.integers(0, 5000))
    if rng.random() < 0.3:
        o.chunk_iters = int(rng.choice([1, 2, 5]))
    return o


def _desc(o):
    return (f"lanes={o.lanes_per_column} lds={o.group_lds} schedule={o.schedule} "
            f"slot={o.slot_elems} pool={o.pool_elems} chunk={o.chunk_iters}")


@pytest.mark.parametrize("seed", range(4))
def test_random_matrices_and_options(seed):
    rng = np.random.default_rng(1000 + seed)
    for i in range(12):
        name, b = _matrix(rng)
        o = _options(rng, b)
        say("random", seed, i, name, b.m, _desc(o))
        ref = oracle.pms(b, snapshot=True)
        low, ess, st = openph_amd.reduce(b, "pms", o)
        assert np.array_equal(low, ref.low) and np.array_equal(ess, ref.ess)
        assert st.additions == ref.additions
        assert st.merge_elems == ref.merge_elems
        assert st.num_iters == ref.num_iters


@pytest.mark.parametrize("seed", range(2))
def test_random_batches(seed):
    rng = np.random.default_rng(2000 + seed)
    for i in range(4):
        mats = [_matrix(rng)[1] for _ in range(int(rng.integers(2, 16)))]
        o = _options(rng, mats[0])
        if o.lanes_per_column == 1:
            o.lanes_per_column = 8
        o.pool_elems = 0
        say("random batch", seed, i, [b.m for b in mats], _desc(o))
        out, st = openph_amd.pms_batch(mats, o)
        adds = 0
        for (low, ess), b in zip(out, mats):
            ref = oracle.pms(b, snapshot=True)
            adds += ref.additions
            assert np.array_equal(low, ref.low) and np.array_equal(ess, ref.ess)
        assert st.additions == adds
