"""GPU test of openph_pms_device_noalloc (the entry point whose working
buffers live in a static device array): called through the C ABI, its
outputs equal openph_pms_device's, bit for bit, on two calls in a row (the
second on a smaller matrix, so the scratch holds the first call's contents),
on a call that has to compact its pools, and sizes beyond its limits are
refused before any work."""
import ctypes
import os
import sys

import numpy as np
import pytest

import openph_amd

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(__file__), "golden")
sys.path.insert(0, os.path.dirname(__file__))
from test_oracle import _gen  # noqa: E402

MAX_M = 3 << 16
MAX_NNZ = 3 * MAX_M
ESCRATCH = -9


def say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    openph_amd.lib()


def _run(fn, b, opt=None):
    import torch
    r, c = b.coo1()
    dev = torch.device("cuda:0")
    rows = torch.from_numpy(r).to(dev)
    cols = torch.from_numpy(c).to(dev)
    # outputs filled with a value no result has, so that stale contents show
    low = torch.full((b.m,), -7, dtype=torch.int32, device=dev)
    ess = torch.full((b.m,), -7, dtype=torch.int32, device=dev)
    st = openph_amd.Stats()
    o = opt or openph_amd.Options()
    torch.cuda.synchronize()
    rc = fn(rows.data_ptr(), cols.data_ptr(), b.m, rows.numel(), low.data_ptr(), ess.data_ptr(),
            None, ctypes.byref(o), ctypes.byref(st))
    torch.cuda.synchronize()
    return rc, low.cpu().numpy(), ess.cpu().numpy(), st


def test_noalloc_equals_device_entry_on_two_calls_in_a_row():
    L = openph_amd.lib()
    for name in ("torus100", "vr40"):  # the second call is smaller
        say("noalloc", name)
        b = _gen(name)
        f = np.load(os.path.join(GOLD, name + ".npz"))
        rc0, low0, ess0, st0 = _run(L.openph_pms_device, b)
        rc1, low1, ess1, st1 = _run(L.openph_pms_device_noalloc, b)
        assert rc0 == 0 and rc1 == 0
        assert np.array_equal(low1, low0) and np.array_equal(ess1, ess0)
        assert np.array_equal(low1, f["low"] + 1) and np.array_equal(ess1, f["ess"])
        assert st1.num_iters == st0.num_iters and st1.additions == st0.additions
        assert st1.merge_elems == st0.merge_elems
        assert st1.schedule in (1, 2)


def test_noalloc_compacts_within_its_pools():
    say("noalloc compaction")
    L = openph_amd.lib()
    b = _gen("vr40")
    opt = openph_amd.Options()
    opt.slot_elems = 1
    opt.pool_elems = int(np.maximum(np.diff(b.colptr), 1).sum())
    rc0, low0, ess0, _ = _run(L.openph_pms_device, b, opt)
    rc1, low1, ess1, st1 = _run(L.openph_pms_device_noalloc, b, opt)
    assert rc0 == 0 and rc1 == 0 and st1.compactions > 0
    assert np.array_equal(low1, low0) and np.array_equal(ess1, ess0)


def test_noalloc_refuses_sizes_beyond_its_scratch():
    L = openph_amd.lib()
    st = openph_amd.Stats()
    o = openph_amd.Options()
    # refused on the host before any pointer is read
    assert L.openph_pms_device_noalloc(None, None, MAX_M + 1, 0, None, None, None,
                                       ctypes.byref(o), ctypes.byref(st)) == ESCRATCH
    assert L.openph_pms_device_noalloc(None, None, 10, MAX_NNZ + 1, None, None, None,
                                       ctypes.byref(o), ctypes.byref(st)) == ESCRATCH
    assert L.openph_strerror(ESCRATCH).startswith(b"input larger")
