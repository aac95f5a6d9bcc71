"""Writes digest fixtures for the two workloads whose outputs are too large
to store whole: the 512-filtration batched mix and a served request stream.

For every member the expected 0-based low array and ess mask come from the
oracle's pms restatement (snapshot schedule) and are checked, before they are
written, against the oracle's twist on growable columns: an independent
reduction order that must reach the same unique reduced low array and
essential mask. Only SHA-256 digests (16 hex digits) of the two arrays as
little-endian int32 are stored, with a digest of the 1-based COO that guards
the generator.

* batched512.json: complexes.batched_mix(512, seed=0), the batched workload
  (VR 2-skeleta, N ~ U{20..80}, cube / sphere / 5-cluster mix).
* served17.json: the first 17 inputs complexes.served_input draws from
  numpy.random.default_rng(0) (70 % VR 2-skeleta of N ~ U{30..200} uniform
  points, 30 % cubical lower-star grids of side ~ U{16..96}); they include
  a VR with N = 190 and cubical grids of side 85 and 81.

Usage: python tests/golden/make_digests.py  (about a minute)
"""
from __future__ import annotations

import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402
from openph_amd import complexes as C  # noqa: E402


def digest(a) -> str:
    return hashlib.sha256(np.ascontiguousarray(a, dtype="<i4").tobytes()).hexdigest()[:16]


def coo_digest(b) -> str:
    r, c = b.coo1()
    h = hashlib.sha256()
    h.update(np.int64(b.m).tobytes())
    h.update(np.ascontiguousarray(r, dtype="<i4").tobytes())
    h.update(np.ascontiguousarray(c, dtype="<i4").tobytes())
    return h.hexdigest()[:16]


def expected(b) -> dict:
    p = oracle.pms(b, snapshot=True)
    t = oracle.twist_dyn(b)
    assert np.array_equal(p.low, t.low) and np.array_equal(p.ess, t.ess), b.name
    return {"m": b.m, "nnz": b.nnz, "coo": coo_digest(b), "low": digest(p.low),
            "ess": digest(p.ess), "num_iters": p.num_iters}


def main():
    oracle.build()
    mats = C.batched_mix(512, seed=0)
    members = [expected(b) for b in mats]
    with open(os.path.join(HERE, "batched512.json"), "w") as f:
        json.dump({"workload": "complexes.batched_mix(512, seed=0)", "members": members}, f,
                  indent=0)
    print("batched512", sum(x["m"] for x in members), "columns")
    rng = np.random.default_rng(0)
    reqs = []
    for i in range(17):
        kind, x = C.served_input(rng)
        b = C.vr_2skeleton(x) if kind == "points" else C.cubical_lower_star(x)
        e = expected(b)
        e.update({"kind": kind, "size": int(x.shape[0])})
        reqs.append(e)
        print("served", i, kind, x.shape[0], b.m, flush=True)
    with open(os.path.join(HERE, "served17.json"), "w") as f:
        json.dump({"workload": "the first 17 complexes.served_input draws of default_rng(0)",
                   "requests": reqs}, f, indent=0)


if __name__ == "__main__":
    main()
