"""Benchmark: plain pms on the Torus-100 VR 2-skeleton (m = 166,750).

One step = one full pms reduction of the boundary matrix through the
device entry point (openph_pms_device) with the 1-based COO already in HBM,
ending with low/ess in HBM. value = column reductions per second =
m * steps / wall time of the timed steps.

Also reported, on the same JSON line:
* roofline: the Phase II kernel (the column additions; the dominant kernel).
  Algorithmic bytes of one launch = 4 * sum over its additions of
  (|a| + |b| + |a xor b|) + 12 * m (classes, low and arglow read once per
  column); duration from HIP events recorded by the library around every
  Phase II launch on the stream it runs on.
* hbm_alg: B_alg = B_merge + B_io + B_state of the whole reduction
  (B_io = 8 nnz + 8 m, B_state = 36 m * iterations) over device time.
* cpu_baseline: the reference's serial twist on one core (its own
  reduce_col_host, oracle/_ref/libref.so, when that build is present; else
  the oracle's fixed-slot restatement) on the same matrix.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]
"""
import argparse
import json
import os
import statistics
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBPS = 8000.0


def cpu_baseline(b, budget_s: float = 10.0):
    import oracle

    oracle.build()
    kind = "port"
    ref_ok = False
    try:
        ref_ok = oracle.ref_available()
    except OSError:
        ref_ok = False
    # slots per column for the fixed-slot layout: the longest column the
    # reduction produces, over the longest input column, plus one
    probe = oracle.twist(b)
    mx = max(1, int(np.diff(b.colptr).max()))
    cw = -(-max(probe.max_len, mx) // mx) + 1
    times = []
    t_end = time.perf_counter() + budget_s
    while time.perf_counter() < t_end or len(times) < 3:
        t0 = time.perf_counter()
        if ref_ok:
            low, ess = oracle.ref_run("twist", b, cw)
            kind = "reference"
        else:
            low = oracle.twist(b, col_width=cw).low
        times.append(time.perf_counter() - t0)
        if len(times) >= 200:
            break
    med = statistics.median(times)
    return {
        "value": b.m / med,
        "unit": "columns/s",
        "cores": 1,
        "kind": kind,
        "sample": (f"serial twist (slots p={cw * mx}) on the full Torus-100 matrix, "
                   f"{len(times)} runs, median {med * 1e3:.1f} ms; "
                   + ("the reference's reduce_col_host compiled from its sources"
                      if kind == "reference" else "C restatement of the reference's twist")),
        "low_sha_match": None,
    }, low


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; --gpus %d is not "
                                   "supported (independent matrices would shard across "
                                   "GPUs with no communication)" % args.gpus}))
        sys.exit(2)

    import torch

    import openph_amd
    from openph_amd import complexes as C

    openph_amd.lib()
    b = C.torus100()
    r, c = b.coo1()
    dev = torch.device("cuda:0")
    dr = torch.from_numpy(r).to(dev)
    dc = torch.from_numpy(c).to(dev)
    low = torch.empty(b.m, dtype=torch.int32, device=dev)
    ess = torch.empty(b.m, dtype=torch.int32, device=dev)
    stream = torch.cuda.current_stream(dev)
    print(f"bench: torus100 m={b.m} nnz={b.nnz}, warmup {args.warmup}", file=sys.stderr, flush=True)
    for _ in range(args.warmup):
        openph_amd.pms_device(dr, dc, b.m, low, ess, stream=stream)
    torch.cuda.synchronize()
    stats = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        stats.append(openph_amd.pms_device(dr, dc, b.m, low, ess, stream=stream))
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    print(f"bench: {args.steps} steps in {wall:.3f} s", file=sys.stderr, flush=True)

    # parity of the benchmarked result against the committed golden low
    gold = np.load(os.path.join(ROOT, "tests", "golden", "torus100.npz"))
    ok = bool(np.array_equal(low.cpu().numpy(), gold["low"] + 1)
              and np.array_equal(ess.cpu().numpy(), gold["ess"]))

    st = stats[-1]
    m, nnz = b.m, b.nnz
    launches = st.phase2_launches
    p2_bytes = 4.0 * st.merge_elems + 12.0 * m * launches
    p2_ms = statistics.median([s.phase2_ms for s in stats])
    dev_ms = statistics.median([s.device_ms for s in stats])
    achieved = p2_bytes / (p2_ms * 1e-3) / 1e9 if p2_ms > 0 else 0.0
    b_alg = 4.0 * st.merge_elems + 8.0 * nnz + 8.0 * m + 36.0 * m * (st.num_iters - 1)
    value = m * args.steps / wall
    out = {
        "metric": "column reductions/s",
        "value": value,
        "unit": "columns/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32",
        "data": "synthetic",
        "config": {"workload": "pms_torus100_vr2", "m": m, "nnz": nnz,
                   "points": 100, "torus": "R=1 r=0.4 sigma=0.03 seed=0",
                   "algo": "pms", "inputs": "1-based COO resident in HBM"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBPS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBPS, "traffic": None,
                     "kernel": "k_phase2_lane",
                     "bytes_per_launch": p2_bytes / max(1, launches),
                     "avg_launch_ms": p2_ms / max(1, launches)},
        "hbm_alg": {"bytes": b_alg, "device_ms": dev_ms,
                    "GBps": b_alg / (dev_ms * 1e-3) / 1e9 if dev_ms > 0 else 0.0},
        "pms": {"num_iters": st.num_iters, "additions": st.additions,
                "merge_elems": st.merge_elems, "compactions": st.compactions,
                "pool_peak": st.pool_peak, "additions_per_s": st.additions * args.steps / wall,
                "parity_vs_golden": ok},
    }
    if not args.no_cpu:
        cb, cpu_low = cpu_baseline(b)
        cb["low_sha_match"] = bool(np.array_equal(cpu_low, gold["low"]))
        out["cpu_baseline"] = cb
        out["speedup_vs_cpu"] = value / cb["value"]
    print(json.dumps(out))


if __name__ == "__main__":
    main()
