"""GPU tests of the filtrations built on the device: the VR 2-skeleton and
the cubical lower-star matrices must equal, entry for entry, the ones
openph_amd.complexes builds on the host (same order, same values, same
dimensions), including inputs with ties; and a matrix built on the device
reduces to the committed golden result."""
import os
import sys

import numpy as np
import pytest

import openph_amd
from openph_amd import complexes as C

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(__file__), "golden")
sys.path.insert(0, os.path.dirname(__file__))


def say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    openph_amd.lib()


def _same(b, out):
    rows, cols, vals, dims = (x.cpu().numpy() for x in out)
    r, c = b.coo1()
    assert np.array_equal(rows, r) and np.array_equal(cols, c)
    assert np.array_equal(dims, b.dim)
    assert np.array_equal(vals.astype(np.float64), b.value)


def _points(kind, n, seed):
    rng = np.random.default_rng(seed)
    if kind == "cube":
        return C.points_cube(rng, n)
    if kind == "torus":
        return C.points_torus(rng, n)
    if kind == "sphere":
        return C.points_sphere(rng, n)
    if kind == "gmm":
        return C.points_gmm(rng, n)
    if kind == "grid":  # many equal distances: ties broken by dimension and vertex tuple
        g = np.stack(np.meshgrid(np.arange(3), np.arange(3), np.arange(2), indexing="ij"), -1)
        return g.reshape(-1, 3)[:n].astype(np.float64)
    if kind == "dup":  # repeated points: zero-length edges
        p = rng.uniform(size=(n, 3))
        p[1::3] = p[0::3][: len(p[1::3])]
        return p
    raise KeyError(kind)


@pytest.mark.parametrize("kind,n", [("cube", 1), ("cube", 2), ("cube", 3), ("cube", 12),
                                    ("cube", 40), ("torus", 100), ("sphere", 30), ("gmm", 55),
                                    ("grid", 18), ("dup", 20)])
def test_vr_matches_host_builder(kind, n):
    import torch
    say("vr", kind, n)
    pts = _points(kind, n, n)
    b = C.vr_2skeleton(pts)
    _same(b, openph_amd.vr2_device(torch.from_numpy(pts).cuda()))


def test_vr_in_other_dimensions():
    import torch
    rng = np.random.default_rng(5)
    for d in (1, 2, 5, 7):
        pts = rng.uniform(size=(15, d))
        _same(C.vr_2skeleton(pts), openph_amd.vr2_device(torch.from_numpy(pts).cuda()))


@pytest.mark.parametrize("shape,kind", [((1, 1, 1), "smooth"), ((2, 3, 4), "smooth"),
                                        ((8, 8, 8), "smooth"), ((16, 16, 16), "smooth"),
                                        ((24, 20, 16), "smooth"), ((10, 10, 10), "ties"),
                                        ((9, 7, 5), "zeros")])
def test_cubical_matches_host_builder(shape, kind):
    import torch
    say("cubical", shape, kind)
    rng = np.random.default_rng(sum(shape))
    if kind == "smooth":
        f = rng.standard_normal(shape).astype(np.float32)
        if min(shape) > 3:
            from scipy.ndimage import gaussian_filter
            f = gaussian_filter(f, sigma=2.0, mode="reflect", truncate=4.0).astype(np.float32)
    elif kind == "ties":
        f = rng.integers(0, 4, size=shape).astype(np.float32)
    else:  # signed zeros and ties at zero
        f = rng.integers(-1, 2, size=shape).astype(np.float32) * np.float32(0.0)
        f[rng.uniform(size=shape) < 0.3] = np.float32(-0.0)
        f[0, 0, 0] = np.float32(1.5)
    b = C.cubical_lower_star(f)
    _same(b, openph_amd.cubical_device(torch.from_numpy(f).cuda()))


def test_device_built_torus_reduces_to_the_golden_result():
    import torch
    say("torus end to end")
    rng = np.random.default_rng(0)
    pts = C.points_torus(rng, 100)
    rows, cols, vals, dims = openph_amd.vr2_device(torch.from_numpy(pts).cuda())
    m = vals.numel()
    low = torch.empty(m, dtype=torch.int32, device="cuda")
    ess = torch.empty_like(low)
    openph_amd.pms_device(rows, cols, m, low, ess)
    torch.cuda.synchronize()
    f = np.load(os.path.join(GOLD, "torus100.npz"))
    assert np.array_equal(low.cpu().numpy() - 1, f["low"])
    assert np.array_equal(ess.cpu().numpy(), f["ess"])


def test_builder_rejects_bad_arguments():
    import torch
    with pytest.raises(openph_amd.OpenPHError):
        openph_amd.vr2_device(torch.zeros((5, 9), dtype=torch.float64, device="cuda"))
    with pytest.raises(openph_amd.OpenPHError):
        openph_amd.cubical_device(torch.zeros((0, 2, 2), dtype=torch.float32, device="cuda"))
