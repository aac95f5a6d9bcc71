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
  reduce_col_host, oracle/_ref/libref.so, when that build is present and its
  fixed-slot layout fits; else the oracle's restatement on growable columns)
  on the same matrix, with the host's CPU model and core counts.
* dropin: the C symbol openph() host to host with an all-zeros low_true, as
  the reference's callers drive it (trackers filled every iteration).
* cubical64: the same figures (rate, Phase II roofline, CPU twist) on the
  cubical lower-star filtration of a 64^3 grid, the largest single matrix.
* batched / served: the 512-filtration batch and the request streams against
  the plain drop-in; the batched leg carries its own Phase II roofline, each
  served COO stream a whole-job one (its requests' B_alg over the burst's
  span).

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--legs a,b,...] [--no-cpu]
"""
import argparse
import json
import os
import statistics
import sys
import time

# Hardware queues for this process's streams, read once when HIP starts (so
# set before anything initialises it; an environment asking for fewer than
# 16, such as HIP's default of 4, is raised). The served legs run 4 worker
# streams beside the streams the earlier legs created, and streams beyond
# the queues share them, their kernels serialised (served vr_cubical after
# the other legs: 2.9x the plain drop-in with 4 queues, 3.0x with 8, 4.6x
# with 16).
try:
    _hwq = int(os.environ.get("GPU_MAX_HW_QUEUES", "4"))
except ValueError:
    _hwq = 4
if _hwq < 16:
    os.environ["GPU_MAX_HW_QUEUES"] = "16"

import numpy as np  # noqa: E402

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBPS = 8000.0


def host_cpu():
    """CPU model and core counts of the host the baseline runs on."""
    model = None
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        usable = None
    return {"cpu_model": model, "nproc": os.cpu_count(), "usable_cores": usable}


def _timed_runs(fn, budget_s, min_runs=3, max_runs=200):
    times, out = [], None
    t_end = time.perf_counter() + budget_s
    while time.perf_counter() < t_end or len(times) < min_runs:
        t0 = time.perf_counter()
        out = fn()
        times.append(time.perf_counter() - t0)
        if len(times) >= max_runs:
            break
    return times, out


def cpu_baseline(b, name: str, budget_s: float = 10.0):
    """The reference's serial twist on one host core, on the whole matrix.

    With the reference's own reduce_col_host / left_to_right_host compiled
    from its sources (oracle/_ref/libref.so) when that build is present and
    its fixed-slot layout (2 m p ints, p = col_width * longest input column)
    stays under 4 GiB; otherwise the oracle's C restatement of the same twist
    on growable columns. Each run times the column layout and the reduction
    (the reference's create_rows_mp and twist loop), not its per-column
    tracker or event calls."""
    import oracle

    oracle.build()
    try:
        ref_ok = oracle.ref_available()
    except OSError:
        ref_ok = False
    probe = oracle.twist_dyn(b)
    mx = max(1, int(np.diff(b.colptr).max()))
    # slots per column: the longest column the reduction produces, plus one
    cw = -(-max(probe.max_len, mx) // mx) + 1
    slot_bytes = 2 * 4 * (b.m + 1) * cw * mx
    use_ref = ref_ok and slot_bytes <= (4 << 30)
    if use_ref:
        times, res = _timed_runs(lambda: oracle.ref_run("twist", b, cw), budget_s)
        low = res[0]
        kind, how = "reference", (f"the reference's reduce_col_host compiled from its sources, "
                                  f"fixed slots p={cw * mx}")
    else:
        times, res = _timed_runs(lambda: oracle.twist_dyn(b), budget_s)
        low = res.low
        kind = "port"
        how = ("the oracle's C restatement of the reference's twist on growable columns"
               + (f" (the reference's fixed-slot layout, p={cw * mx}, would take "
                  f"{slot_bytes / 2**30:.1f} GiB)" if ref_ok else ""))
    med = statistics.median(times)
    out = {
        "value": b.m / med,
        "unit": "columns/s",
        "cores": 1,
        "kind": kind,
        "sample": (f"serial twist on the full {name} matrix (m={b.m}), {len(times)} runs of "
                   f"about {budget_s:.0f} s, median {med * 1e3:.1f} ms; {how}"),
        "low_sha_match": None,
    }
    out.update(host_cpu())
    return out, low


def device_leg(b, name, steps, warmup, gold=None, timed_steps=5):
    """Full pms reductions of `b` through openph_pms_device, COO in HBM
    before timing starts, low/ess left in HBM, in two passes:

    * `steps` timed reductions with the automatic schedule (iterations
      queued in chunks, replayed from a cached hipGraph once the size
      repeats) and no per-launch markers: the rate;
    * `timed_steps` more on the queued schedule with HIP events on the
      stream around every Phase II launch (options.time_phase2): that
      kernel's launch times, for its roofline.

    Returns the wall time of the first pass, the stats of both and whether
    the result equals `gold`."""
    import torch

    import openph_amd

    r, c = b.coo1()
    dev = torch.device("cuda:0")
    dr = torch.from_numpy(r).to(dev)
    dc = torch.from_numpy(c).to(dev)
    low = torch.empty(b.m, dtype=torch.int32, device=dev)
    ess = torch.empty(b.m, dtype=torch.int32, device=dev)
    stream = torch.cuda.current_stream(dev)
    opt = openph_amd.Options()
    print(f"bench: {name} m={b.m} nnz={b.nnz}, warmup {warmup}, steps {steps}",
          file=sys.stderr, flush=True)
    for _ in range(warmup):
        openph_amd.pms_device(dr, dc, b.m, low, ess, stream=stream, options=opt)
    torch.cuda.synchronize()
    stats = []
    t0 = time.perf_counter()
    for _ in range(steps):
        stats.append(openph_amd.pms_device(dr, dc, b.m, low, ess, stream=stream, options=opt))
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    print(f"bench: {name}: {steps} steps in {wall:.3f} s", file=sys.stderr, flush=True)
    ok = None
    if gold is not None:
        ok = bool(np.array_equal(low.cpu().numpy(), gold[0] + 1)
                  and np.array_equal(ess.cpu().numpy(), gold[1]))
    topt = openph_amd.Options()
    topt.schedule = openph_amd.Options.QUEUED
    topt.time_phase2 = 1
    tstats = [openph_amd.pms_device(dr, dc, b.m, low, ess, stream=stream, options=topt)
              for _ in range(timed_steps)]
    torch.cuda.synchronize()
    return wall, stats, tstats, ok


def phase2_roofline(stats, m, kernel):
    """Roofline of the Phase II launches: algorithmic bytes per launch =
    (4 * sum over additions of (|a| + |b| + |a xor b|) + 12 * m * launches)
    / launches (each launch reads low, arglow and classes of every column
    once), over the average launch time (median over the reductions of
    `stats`, which were run with options.time_phase2)."""
    st = stats[-1]
    launches = max(1, st.phase2_launches)
    p2_bytes = 4.0 * st.merge_elems + 12.0 * m * launches
    p2_ms = statistics.median([s.phase2_ms for s in stats])
    achieved = p2_bytes / (p2_ms * 1e-3) / 1e9 if p2_ms > 0 else 0.0
    return {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBPS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBPS, "traffic": None, "kernel": kernel,
            "bytes_per_launch": p2_bytes / launches, "avg_launch_ms": p2_ms / launches,
            "launches": st.phase2_launches}


def b_alg(st, m, nnz):
    """B_merge + B_io + B_state of one reduction (36 m bytes per iteration)."""
    return 4.0 * st.merge_elems + 8.0 * nnz + 8.0 * m + 36.0 * m * (st.num_iters - 1)


def cubical_leg(steps: int = 8, warmup: int = 2, cpu: bool = True):
    """Plain pms on the cubical lower-star filtration of a 64^3 Gaussian field
    (sigma 2, seed 0; m = 2,048,383), the largest single-matrix
    configuration: column reductions/s, its Phase II roofline and the CPU
    twist on the same matrix."""
    from openph_amd import complexes as C

    t0 = time.perf_counter()
    b = C.cubical_lower_star(C.smoothed_field(np.random.default_rng(0), 64))
    gen_s = time.perf_counter() - t0
    wall, stats, tstats, _ = device_leg(b, "cubical64", steps, warmup)
    st = tstats[-1]
    dev_ms = statistics.median([s.device_ms for s in stats])
    value = b.m * steps / wall
    roof = phase2_roofline(tstats, b.m, f"Phase II: k_phase2_group<{st.lanes}, {st.group_lds}> "
                                       "+ k_phase2_wide + k_phase2_long")
    # HBM bytes per Phase II pass from the committed counter runs
    tfile = os.path.join(ROOT, "profiles", "pms_cubical64_traffic.json")
    if os.path.exists(tfile):
        tj = json.load(open(tfile))
        if f"k_phase2_group<{st.lanes}, {st.group_lds}>" in tj.get("kernel", ""):
            roof["traffic"] = tj["traffic_bytes_per_launch"]
    out = {
        "workload": "pms_cubical64_lower_star", "m": b.m, "nnz": b.nnz,
        "field": "64^3 N(0,1) float32 seed 0, Gaussian smoothing sigma 2",
        "cols_per_s": value, "ms_per_step": wall / steps * 1e3, "steps": steps,
        "roofline": roof,
        "hbm_alg": {"bytes": b_alg(st, b.m, b.nnz), "device_ms": dev_ms,
                    "GBps": b_alg(st, b.m, b.nnz) / (dev_ms * 1e-3) / 1e9},
        "pms": {"num_iters": st.num_iters, "additions": st.additions,
                "merge_elems": st.merge_elems, "phase2_ms": statistics.median(
                    [s.phase2_ms for s in tstats]), "handoffs": st.handoffs,
                "schedule": stats[-1].schedule,
                "compactions": st.compactions, "pool_peak": st.pool_peak},
        "host_generation_s": gen_s,
    }
    if cpu:
        print("bench: cubical64 cpu baseline", file=sys.stderr, flush=True)
        cb, cpu_low = cpu_baseline(b, "cubical 64^3")
        out["cpu_baseline"] = cb
        out["speedup_vs_cpu"] = value / cb["value"]
        out["cpu_equals_gpu"] = None  # filled by main from the GPU result
        out["_cpu_low"] = cpu_low
        out["_b"] = b
        out["reference_sample"] = cubical_reference_sample()
    return out


def cubical_reference_sample(side: int = 40, budget_s: float = 6.0):
    """The reference's own serial twist (reduce_col_host compiled from its
    sources, fixed slots) on a bounded sample of the cubical workload: the
    same field recipe on a side^3 grid, whose slot layout still fits 4 GiB
    (64^3 would need about 28 GiB), beside the restatement and pms on the
    same matrix. On cubical grids the reference's fixed slots (p = width x
    longest input column, with the width the reduction's longest column
    needs) dominate its time, so the restatement that cubical64's
    cpu_baseline times is the stronger baseline."""
    import oracle
    from openph_amd import complexes as C

    try:
        oracle.build()
        if not oracle.ref_available():
            return None
    except OSError:
        return None
    b = C.cubical_lower_star(C.smoothed_field(np.random.default_rng(0), side))
    probe = oracle.twist_dyn(b)
    mx = max(1, int(np.diff(b.colptr).max()))
    cw = -(-max(probe.max_len, mx) // mx) + 1
    slot_bytes = 2 * 4 * (b.m + 1) * cw * mx
    if slot_bytes > (4 << 30):
        return None
    print(f"bench: cubical {side}^3 reference sample", file=sys.stderr, flush=True)
    t_ref, res = _timed_runs(lambda: oracle.ref_run("twist", b, cw), budget_s)
    t_port, _ = _timed_runs(lambda: oracle.twist_dyn(b), budget_s / 3)
    wall, stats, _, _ = device_leg(b, f"cubical{side}", 10, 2, timed_steps=1)
    ref_ms, port_ms = statistics.median(t_ref) * 1e3, statistics.median(t_port) * 1e3
    gpu_ms = wall / 10 * 1e3
    return {"workload": f"cubical {side}^3 lower star, same field recipe (m={b.m})",
            "kind": "reference", "cores": 1,
            "reference_ms": ref_ms, "reference_runs": len(t_ref), "slots_p": cw * mx,
            "slot_bytes": slot_bytes, "restatement_ms": port_ms, "gpu_ms": gpu_ms,
            "gpu_vs_reference": ref_ms / gpu_ms, "gpu_vs_restatement": port_ms / gpu_ms,
            "reference_low_equals_restatement": bool(np.array_equal(res[0], probe.low))}


def vr40_leg(steps: int = 50, warmup: int = 5, cpu: bool = True):
    """Plain pms on VR-40 (40 points i.i.d. U[0,1]^3, seed 0; m = 10,700),
    the reference's CPU-sized configuration: column reductions/s with the
    COO in HBM and host to host (openph_v2: host COO in, host low/ess out),
    its Phase II roofline, B_alg over device time, and the CPU twist on the
    same matrix."""
    import openph_amd
    from openph_amd import complexes as C

    b = C.vr40()
    gold = np.load(os.path.join(ROOT, "tests", "golden", "vr40.npz"))
    wall, stats, tstats, ok = device_leg(b, "vr40", steps, warmup, (gold["low"], gold["ess"]))
    st = tstats[-1]
    dev_ms = statistics.median([s.device_ms for s in stats])
    value = b.m * steps / wall
    kname = (f"k_phase2_group<{st.lanes}, {st.group_lds}>" if st.lanes > 1 else "k_phase2_lane")
    openph_amd.reduce(b, "pms")  # host-to-host warm-up
    ts = []
    for _ in range(20):
        t0 = time.perf_counter()
        openph_amd.reduce(b, "pms")
        ts.append(time.perf_counter() - t0)
    host_s = statistics.median(ts)
    out = {
        "workload": "pms_vr40_cube", "m": b.m, "nnz": b.nnz,
        "points": "40 i.i.d. U[0,1]^3, seed 0",
        "cols_per_s": value, "ms_per_step": wall / steps * 1e3, "steps": steps,
        "host_to_host_cols_per_s": b.m / host_s, "host_to_host_ms": host_s * 1e3,
        "roofline": phase2_roofline(tstats, b.m, kname),
        "hbm_alg": {"bytes": b_alg(st, b.m, b.nnz), "device_ms": dev_ms,
                    "GBps": b_alg(st, b.m, b.nnz) / (dev_ms * 1e-3) / 1e9 if dev_ms > 0 else 0.0},
        "pms": {"num_iters": st.num_iters, "additions": st.additions,
                "merge_elems": st.merge_elems, "schedule": stats[-1].schedule,
                "parity_vs_golden": ok},
    }
    if cpu:
        print("bench: vr40 cpu baseline", file=sys.stderr, flush=True)
        cb, cpu_low = cpu_baseline(b, "VR-40", budget_s=3.0)
        cb["low_sha_match"] = bool(np.array_equal(cpu_low, gold["low"]))
        out["cpu_baseline"] = cb
        out["speedup_vs_cpu"] = value / cb["value"]
    return out


def dropin_leg(b, reps: int = 20):
    """The drop-in as the reference's callers drive it: the C symbol
    openph() (src/cuda/openph.cu:29-35) host to host on Torus-100 with an
    all-zeros low_true (numerics/benchmark_datasets.m:53), so the trackers
    are filled every iteration, and the outputs initialised as the mex
    gateway does (src/matlab/ph/ph.cu:72-78)."""
    import ctypes

    import openph_amd

    L = openph_amd.lib()
    r, c = b.coo1()
    m = b.m
    lt = np.zeros(m, np.int32)
    low = np.full(m, -1, np.int32)
    ess = np.full(m, -1, np.int32)
    errs = [np.full(m + 1, -1, np.float32) for _ in range(4)]
    tt = np.zeros(m + 1, np.float32)
    nit = np.zeros(1, np.int32)
    ip = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))  # noqa: E731
    fp = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))  # noqa: E731

    def call():
        L.openph(b"pms", ip(r), ip(c), m, 8, ip(lt), r.shape[0], ip(low), ip(ess), fp(errs[0]),
                 fp(errs[1]), fp(errs[2]), fp(errs[3]), fp(tt), ip(nit), 0)

    for _ in range(3):
        call()
    ts = []
    for _ in range(reps):
        t0 = time.perf_counter()
        call()
        ts.append(time.perf_counter() - t0)
    # the schedule the drop-in picks with trackers on (openph_v2, same options)
    opt = openph_amd.Options()
    opt.track = 1
    st = openph_amd.Stats()
    trk = np.zeros((5, m + 1), np.float32)
    L.openph_v2(b"pms", ip(r), ip(c), m, r.shape[0], ip(lt), ctypes.byref(opt), ip(low), ip(ess),
                fp(trk), ctypes.byref(st))
    med = statistics.median(ts)
    return {"workload": "openph('pms') on Torus-100, host COO in, host outputs, low_true = 0",
            "ms_per_call": med * 1e3, "cols_per_s": m / med, "calls": reps,
            "num_iters": int(nit[0]), "schedule": st.schedule, "host_reads": st.host_reads,
            "time_track_sum_ms": float(tt[1:int(nit[0])].sum())}


def batched_leg(count: int = 512, reps: int = 3, tune: bool = True):
    """512 VR 2-skeleta (N ~ U{20..80}, mixed shapes) reduced (a) by one
    batched call with the setting the autotuner picks on this batch and
    (b) by `count` sequential plain calls with the default setting, inputs
    resident in HBM for both; outputs compared bit for bit."""
    import torch

    import openph_amd
    from openph_amd import complexes as C

    t0 = time.perf_counter()
    mats = C.batched_mix(count, seed=0)
    gen_s = time.perf_counter() - t0
    ms = np.array([b.m for b in mats], np.int32)
    nnzs = np.array([b.nnz for b in mats], np.int32)
    rr, cc = zip(*(b.coo1() for b in mats))
    dev = torch.device("cuda:0")
    rows = torch.from_numpy(np.concatenate(rr)).to(dev)
    cols = torch.from_numpy(np.concatenate(cc)).to(dev)
    M = int(ms.sum())
    low_b = torch.empty(M, dtype=torch.int32, device=dev)
    ess_b = torch.empty(M, dtype=torch.int32, device=dev)
    low_p = torch.empty(M, dtype=torch.int32, device=dev)
    ess_p = torch.empty(M, dtype=torch.int32, device=dev)
    stream = torch.cuda.current_stream(dev)
    moff = np.concatenate([[0], np.cumsum(ms)]).astype(np.int64)
    noff = np.concatenate([[0], np.cumsum(nnzs)]).astype(np.int64)

    def plain_pass():
        for k in range(count):
            a, b_ = int(noff[k]), int(noff[k + 1])
            x, y = int(moff[k]), int(moff[k + 1])
            openph_amd.pms_device(rows[a:b_], cols[a:b_], y - x, low_p[x:y], ess_p[x:y],
                                  stream=stream)

    best, tune_rows, tune_s = None, [], 0.0
    if tune:
        t0 = time.perf_counter()
        best, tune_rows = openph_amd.autotune_batch_device(ms, nnzs, rows, cols, low_b, ess_b,
                                                           stream=stream, reps=2)
        tune_s = time.perf_counter() - t0
    # warm both paths once
    st_b = openph_amd.pms_batch_device(ms, nnzs, rows, cols, low_b, ess_b, stream=stream,
                                       options=best)
    plain_pass()
    torch.cuda.synchronize()
    tb = []
    for _ in range(reps):
        t0 = time.perf_counter()
        st_b = openph_amd.pms_batch_device(ms, nnzs, rows, cols, low_b, ess_b, stream=stream,
                                           options=best)
        torch.cuda.synchronize()
        tb.append(time.perf_counter() - t0)
    t0 = time.perf_counter()
    plain_pass()
    torch.cuda.synchronize()
    tp = time.perf_counter() - t0
    same = bool(torch.equal(low_b, low_p) and torch.equal(ess_b, ess_p))
    # one more batched call with every Phase II launch timed, for the roofline
    topt = (openph_amd.Options.from_buffer_copy(best) if best is not None
            else openph_amd.Options())
    topt.time_phase2 = 1
    st_t = openph_amd.pms_batch_device(ms, nnzs, rows, cols, low_b, ess_b, stream=stream,
                                       options=topt)
    torch.cuda.synchronize()
    b_s = statistics.median(tb)
    return {
        "workload": f"{count} VR 2-skeleta, N~U{{20..80}}, cube/sphere/5-GMM, seed 0",
        "columns": M, "nnz": int(nnzs.sum()),
        "batched_cols_per_s": M / b_s, "batched_ms": b_s * 1e3,
        "plain_sequential_cols_per_s": M / tp, "plain_sequential_ms": tp * 1e3,
        "speedup_batched_vs_plain": tp / b_s,
        "tuned": ({"lanes_per_column": st_b.lanes, "group_lds": st_b.group_lds,
                   "slot_elems": best.slot_elems or 32, "schedule": st_b.schedule,
                   "tuning_s": tune_s, "candidates": len(tune_rows),
                   "all_candidates_equal": all(r["same"] for r in tune_rows)}
                  if tune else None),
        "batched_num_iters": st_b.num_iters, "batched_additions": st_b.additions,
        "batched_merge_elems": st_b.merge_elems,
        "roofline": dict(phase2_roofline([st_t], M, f"k_phase2_group<{st_t.lanes}, "
                                                     f"{st_t.group_lds}> + k_phase2_mid"),
                         timing=("device-clock stamps around each Phase II launch "
                                 "(graph schedule), one more batched call"
                                 if st_t.schedule == 3 else
                                 "HIP events around each Phase II launch, one more batched "
                                 "call")),
        "batched_equals_plain": same, "host_generation_s": gen_s,
    }


def gpu_twist_leg(b, reps: int = 3):
    """Torus-100 reduced by the twist algorithm on the same GPU (the
    reference's column order, one workgroup taking the columns in turn) and
    by pms, both through the host entry point (host COO in, host low/ess
    out); outputs compared. Twist's 3 M dependent column additions take
    seconds on one workgroup, so it runs once."""
    import openph_amd

    out = {}
    res = {}
    for algo in ("twist", "pms"):
        n = 1 if algo == "twist" else reps
        if algo == "pms":
            openph_amd.reduce(b, algo)  # warm-up
        ts = []
        for _ in range(n):
            t0 = time.perf_counter()
            low, ess, st = openph_amd.reduce(b, algo)
            ts.append(time.perf_counter() - t0)
        res[algo] = (low, ess)
        out[f"{algo}_ms"] = statistics.median(ts) * 1e3
        out[f"{algo}_cols_per_s"] = b.m / statistics.median(ts)
        out[f"{algo}_additions"] = st.additions
    out["pms_vs_twist"] = out["twist_ms"] / out["pms_ms"]
    out["same_result"] = bool(np.array_equal(res["twist"][0], res["pms"][0])
                              and np.array_equal(res["twist"][1], res["pms"][1]))
    return out


def builder_leg(reps: int = 5):
    """The boundary matrices built on the device (openph_vr2_device,
    openph_cubical_device) against the host builder (openph_amd.complexes),
    from the same point cloud / field; outputs compared."""
    import torch

    import openph_amd
    from openph_amd import complexes as C

    out = {}
    rng = np.random.default_rng(0)
    pts = C.points_torus(rng, 100)
    field = C.smoothed_field(np.random.default_rng(0), 64)
    for name, host, dev_in, build in (
            ("vr_torus100", lambda: C.vr_2skeleton(pts), torch.from_numpy(pts).cuda(),
             openph_amd.vr2_device),
            ("cubical64", lambda: C.cubical_lower_star(field), torch.from_numpy(field).cuda(),
             openph_amd.cubical_device)):
        t0 = time.perf_counter()
        b = host()
        t_host = time.perf_counter() - t0
        res = build(dev_in)
        torch.cuda.synchronize()
        ts = []
        for _ in range(reps):
            t0 = time.perf_counter()
            res = build(dev_in)
            torch.cuda.synchronize()
            ts.append(time.perf_counter() - t0)
        r, c = b.coo1()
        same = bool(np.array_equal(res[0].cpu().numpy(), r) and np.array_equal(res[1].cpu().numpy(), c))
        out[name] = {"m": b.m, "nnz": b.nnz, "device_ms": statistics.median(ts) * 1e3,
                     "host_ms": t_host * 1e3, "speedup": t_host / statistics.median(ts),
                     "same_matrix": same}
    return out


def served_leg(workers: int = 4, seed: int = 0):
    """Served pms (openph_server_*: worker streams, admission by pool
    footprint, continuous batching) against the plain drop-in (openph_v2,
    host COO in, host low/ess out, one request after another), on two
    request streams:

    * mix512: the 512 VR 2-skeleta of the batched leg as 512 requests;
    * vr_cubical: 72 requests cycling over 24 drawn once (seed 0): 70% VR
      2-skeleta of N ~ U{30..200} uniform points, 30% cubical lower-star
      grids of side ~ U{16..96} (Gaussian field smoothed with sigma 2).

    For each: throughput with every request submitted at once (the server
    saturated; sum of m / (last result - first submit)), and p50/p99 latency
    (submit -> results on the host) under Poisson arrivals at the rate the
    plain drop-in sustains. Every result is compared with the same matrix
    reduced alone."""
    from openph_amd import complexes as C

    t0 = time.perf_counter()
    mix = C.batched_mix(512, seed=0)
    rng = np.random.default_rng(seed)
    pool = [C.served_request(rng) for _ in range(24)]
    gen_s = time.perf_counter() - t0
    out = {"workers": workers, "host_generation_s": gen_s}
    out["mix512"] = _served_stream(mix, list(range(len(mix))), workers, seed,
                                   "the 512 VR 2-skeleta of the batched leg")
    out["vr_cubical"] = _served_stream(
        pool, [i % len(pool) for i in range(72)], workers, seed,
        "72 requests cycling over 24 (seed 0): 70% VR N~U{30..200}, 30% cubical side~U{16..96}")
    out["vr_cubical_built"] = _served_built_stream(
        [i % 24 for i in range(72)], workers, seed,
        "the vr_cubical stream given as point clouds and fields, built on the device")
    return out


def _served_built_stream(seq, workers, seed, what):
    """Requests given as raw inputs (points / fields). Plain: each request
    uploaded, built (openph_vr2_device / openph_cubical_device) and reduced
    (openph_pms_device) alone, results copied back, one after another.
    Served: the server builds each request on its worker's stream inside the
    batch."""
    import threading

    import torch

    import openph_amd
    from openph_amd import complexes as C
    from openph_amd.server import Server

    rng = np.random.default_rng(seed)
    inputs = [C.served_input(rng) for _ in range(24)]

    def plain_one(kind, x):
        t = torch.from_numpy(x).cuda()
        rows, cols, vals, _ = (openph_amd.vr2_device(t) if kind == "points"
                               else openph_amd.cubical_device(t))
        m = vals.numel()
        low = torch.empty(m, dtype=torch.int32, device="cuda")
        ess = torch.empty_like(low)
        openph_amd.pms_device(rows, cols, m, low, ess)
        return low.cpu().numpy() - 1, ess.cpu().numpy()

    ref = [plain_one(k, x) for k, x in inputs]  # warm-up and outputs alone
    cols = sum(ref[i][0].shape[0] for i in seq)
    t0 = time.perf_counter()
    for i in seq:
        plain_one(*inputs[i])
    t_plain = time.perf_counter() - t0

    def run(qps, rng2):
        n_requests = len(seq)
        arrivals = (np.zeros(n_requests) if qps is None
                    else np.cumsum(rng2.exponential(1.0 / qps, n_requests)))
        tickets = [None] * n_requests
        with Server(workers=workers) as srv:
            def submitter():
                start = time.perf_counter()
                for k, i in enumerate(seq):
                    dt = arrivals[k] - (time.perf_counter() - start)
                    if dt > 0:
                        time.sleep(dt)
                    kind, x = inputs[i]
                    tickets[k] = (srv.submit_points(x, borrow=True) if kind == "points"
                                  else srv.submit_field(x, borrow=True))

            th = threading.Thread(target=submitter)
            th.start()
            lat, first, last, same = [], None, None, True
            for k, i in enumerate(seq):
                while tickets[k] is None:
                    time.sleep(0.0002)
                low, ess, info = srv.wait(tickets[k])
                same &= bool(np.array_equal(low, ref[i][0]) and np.array_equal(ess, ref[i][1]))
                lat.append(info.done_s - info.submit_s)
                first = info.submit_s if first is None else min(first, info.submit_s)
                last = info.done_s if last is None else max(last, info.done_s)
            th.join()
        lat_ms = np.array(lat) * 1e3
        return {"qps": qps, "served_cols_per_s": cols / (last - first), "span_s": last - first,
                "latency_ms_p50": float(np.percentile(lat_ms, 50)),
                "latency_ms_p99": float(np.percentile(lat_ms, 99)), "served_equals_alone": same}

    burst = run(None, np.random.default_rng(seed + 1))
    paced = run(len(seq) / t_plain, np.random.default_rng(seed + 1))
    return {
        "workload": what, "requests": len(seq), "columns": cols,
        "plain_build_and_reduce_cols_per_s": cols / t_plain, "plain_s": t_plain,
        "served_cols_per_s": burst["served_cols_per_s"],
        "speedup_served_vs_plain": burst["served_cols_per_s"] / (cols / t_plain),
        "latency_ms_p50": paced["latency_ms_p50"], "latency_ms_p99": paced["latency_ms_p99"],
        "burst": burst, "poisson": paced,
        "served_equals_alone": burst["served_equals_alone"] and paced["served_equals_alone"],
    }


def _served_stream(pool, seq, workers, seed, what):
    import openph_amd

    coo = [b.coo1() for b in pool]
    alone = [openph_amd.reduce(b, "pms") for b in pool]  # warm-up and outputs alone
    ref = [a[:2] for a in alone]
    # algorithmic bytes of each request, from its counters reduced alone
    # (a batch member performs the same additions)
    balg = [b_alg(a[2], b.m, b.nnz) for a, b in zip(alone, pool)]
    cols = sum(pool[i].m for i in seq)
    t0 = time.perf_counter()
    dev_ms = 0.0
    for i in seq:
        dev_ms += openph_amd.reduce(pool[i], "pms")[2].device_ms
    t_plain = time.perf_counter() - t0
    # an untimed pass first: the server path's first use in the process
    _served_run(pool, coo, ref, seq, cols, None, workers, np.random.default_rng(seed + 1))
    burst = _served_run(pool, coo, ref, seq, cols, None, workers, np.random.default_rng(seed + 1))
    copied = _served_run(pool, coo, ref, seq, cols, None, workers, np.random.default_rng(seed + 1),
                         borrow=False)
    qps = len(seq) / t_plain
    paced = _served_run(pool, coo, ref, seq, cols, qps, workers, np.random.default_rng(seed + 1))
    job_bytes = float(sum(balg[i] for i in seq))
    achieved = job_bytes / burst["span_s"] / 1e9
    return {
        "workload": what, "requests": len(seq), "columns": cols,
        # whole job: the requests' B_alg (merges, COO in, outputs, 36 m per
        # iteration) over the burst's span, every kernel of every batch
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBPS, "traffic": None,
                     "kernel": "whole served job (all kernels of all batches)",
                     "bytes": job_bytes, "span_s": burst["span_s"]},
        "plain_sequential_cols_per_s": cols / t_plain, "plain_sequential_s": t_plain,
        # the device time of those reductions (set-up to outputs, events on
        # the drop-in's stream): the rest of plain_sequential_s is host work
        "plain_device_s": dev_ms / 1e3,
        "served_cols_per_s": burst["served_cols_per_s"],
        "speedup_served_vs_plain": burst["served_cols_per_s"] / (cols / t_plain),
        "latency_ms_p50": paced["latency_ms_p50"], "latency_ms_p99": paced["latency_ms_p99"],
        "latency_qps": qps, "burst": burst, "poisson": paced,
        "burst_copied_submit": {k: copied[k] for k in
                                ("served_cols_per_s", "span_s", "submit_span_s")},
        "served_equals_alone": (burst["served_equals_alone"] and paced["served_equals_alone"]
                                and copied["served_equals_alone"]),
    }


def _served_run(pool, coo, ref, seq, cols, qps, workers, rng, borrow=True):
    """Submit `seq` to a fresh server, all at once (qps None) or with Poisson
    arrivals at `qps`; wait for every result and compare it with `ref`.
    borrow: zero-copy submission (the worker reads the caller's COO in place
    and writes the results into arrays handed over at submit, as the
    reference's mex gateway lends its inputs and outputs); otherwise each
    request is copied at submit."""
    import threading

    from openph_amd.server import Server

    n_requests = len(seq)
    arrivals = (np.zeros(n_requests) if qps is None
                else np.cumsum(rng.exponential(1.0 / qps, n_requests)))
    tickets = [None] * n_requests
    kinds = []
    with Server(workers=workers) as srv:
        def submitter():
            start = time.perf_counter()
            for k, i in enumerate(seq):
                dt = arrivals[k] - (time.perf_counter() - start)
                if dt > 0:
                    time.sleep(dt)
                r, c = coo[i]
                tickets[k] = srv.submit_coo(r, c, pool[i].m, borrow=borrow)

        th = threading.Thread(target=submitter)
        th.start()
        lat, first, last, same = [], None, None, True
        last_submit = None
        for k, i in enumerate(seq):
            while tickets[k] is None:
                time.sleep(0.0002)
            low, ess, info = srv.wait(tickets[k])
            same &= bool(np.array_equal(low, ref[i][0]) and np.array_equal(ess, ref[i][1]))
            lat.append(info.done_s - info.submit_s)
            kinds.append(pool[i].m)
            first = info.submit_s if first is None else min(first, info.submit_s)
            last = info.done_s if last is None else max(last, info.done_s)
            last_submit = info.submit_s if last_submit is None else max(last_submit, info.submit_s)
        th.join()
        st = srv.stats()
    span = last - first
    lat_ms = np.array(lat) * 1e3
    # latency per request class: the VR 2-skeleta (input columns of at most
    # 3 entries) and the cubical grids
    cls = np.array([int(np.diff(pool[i].colptr).max(initial=0) > 4) for i in seq])
    by = {}
    for c, name in ((0, "vr"), (1, "cubical")):
        x = lat_ms[cls == c]
        if x.size:
            by[name] = {"requests": int(x.size), "latency_ms_p50": float(np.percentile(x, 50)),
                        "latency_ms_p99": float(np.percentile(x, 99))}
    return {
        "qps": qps, "served_cols_per_s": cols / span, "span_s": span,
        "submit_span_s": last_submit - first,
        "latency_ms_p50": float(np.percentile(lat_ms, 50)),
        "latency_ms_p99": float(np.percentile(lat_ms, 99)),
        "latency_ms_max": float(lat_ms.max()), "by_class": by,
        "batches": st["batches"], "served_equals_alone": same,
        "submit": "borrowed" if borrow else "copied",
    }


LEGS = ("torus", "vr40", "dropin", "cubical", "twist", "builders", "batched", "served", "cpu",
        "batched_only", "phase_boundaries", "iterations")
DEFAULT_LEGS = ("torus", "vr40", "dropin", "cubical", "twist", "builders", "batched", "served",
                "cpu")


def iterations_leg(calls: int = 5):
    """Per-iteration device time of Torus-100 and cubical 64^3 as the
    reference's time_track reports it: ph('pms') with an all-zeros low_true
    (the trackers on, as numerics/benchmark_datasets.m:53 calls it), each
    iteration timed on the device clock from its first kernel to its last
    (pms.cu:46-103). Median over `calls` calls, per iteration, in us."""
    import openph_amd
    from openph_amd import complexes as C

    out = {}
    for name, b in (("torus100", C.torus100()),
                    ("cubical64", C.cubical_lower_star(C.smoothed_field(
                        np.random.default_rng(0), 64)))):
        r, c = b.coo1()
        lt = np.zeros(b.m, np.int32)
        runs = []
        for _ in range(calls + 1):  # the first call untimed
            o = openph_amd.ph("pms", r, c, b.m, 8, lt)
            n = int(o[7][0])
            runs.append(o[6][1:n].astype(np.float64) * 1e3)
        t = np.median(np.array(runs[1:]), axis=0)
        out[name] = {"m": b.m, "num_iters": n, "iteration_us": [round(float(x), 1) for x in t],
                     "sum_us": float(t.sum())}
    return out


def phase_boundaries_leg(phases: int = 6, reps: int = 200):
    """What a persistent (cooperative) iteration would trade: the same chain
    of `phases` grid-wide passes (each reads what the previous wrote across
    the grid) separated by kernel boundaries, queued or replayed from a
    hipGraph, or by grid barriers inside one cooperative launch (one arrival
    counter, or one per XCD group of blocks and then one for the groups), at
    the column counts of Torus-100 and cubical 64^3. libopenph_probe.so
    (openph_amd/csrc/phase_probe.hip); every form must leave the same array."""
    import ctypes

    L = ctypes.CDLL(os.path.join(ROOT, "openph_amd", "libopenph_probe.so"))
    L.openph_probe_phases.restype = ctypes.c_int
    L.openph_probe_phases.argtypes = [ctypes.c_int] * 5 + [
        ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_int),
        ctypes.POINTER(ctypes.c_longlong)]
    out = {"phases": phases, "reps": reps,
           "pass": "y[j] = x[j] + x[(j + m/2) mod m], 4 columns per thread"}
    import torch

    cus = torch.cuda.get_device_properties(0).multi_processor_count
    forms = (("queued", 0, 0), ("graph", 1, 0), ("cooperative_flat", 2, 0),
             ("cooperative_two_level", 3, 0), ("cooperative_flat_one_block_per_cu", 2, cus),
             ("cooperative_two_level_one_block_per_cu", 3, cus))
    for label, m in (("torus100_m", 166750), ("cubical64_m", 2048383)):
        row, sums = {"m": m}, set()
        for name, mode, cap in forms:
            ms, blocks, cs = ctypes.c_float(), ctypes.c_int(), ctypes.c_longlong()
            rc = L.openph_probe_phases(m, phases, reps, mode, cap, ctypes.byref(ms),
                                       ctypes.byref(blocks), ctypes.byref(cs))
            if rc != 0:
                row[name] = {"status": rc}
                continue
            sums.add(cs.value)
            row[name] = {"us_per_chain": ms.value * 1e3 / reps,
                         "us_per_pass": ms.value * 1e3 / reps / phases, "blocks": blocks.value}
        row["same_result"] = len(sums) == 1
        out[label] = row
    return out


def batched_only_leg(count: int = 512, reps: int = 5):
    """The batched call alone (default options, no tuning and no plain
    pass): the workload of the batched kernel profile."""
    import torch

    import openph_amd
    from openph_amd import complexes as C

    mats = C.batched_mix(count, seed=0)
    ms = np.array([b.m for b in mats], np.int32)
    nnzs = np.array([b.nnz for b in mats], np.int32)
    rr, cc = zip(*(b.coo1() for b in mats))
    dev = torch.device("cuda:0")
    rows = torch.from_numpy(np.concatenate(rr)).to(dev)
    cols = torch.from_numpy(np.concatenate(cc)).to(dev)
    M = int(ms.sum())
    low = torch.empty(M, dtype=torch.int32, device=dev)
    ess = torch.empty(M, dtype=torch.int32, device=dev)
    stream = torch.cuda.current_stream(dev)
    opt = openph_amd.Options()
    opt.schedule = openph_amd.Options.QUEUED
    opt.time_phase2 = 1
    stats, ts = [], []
    for _ in range(reps):
        t0 = time.perf_counter()
        stats.append(openph_amd.pms_batch_device(ms, nnzs, rows, cols, low, ess, stream=stream,
                                                 options=opt))
        torch.cuda.synchronize()
        ts.append(time.perf_counter() - t0)
    st = stats[-1]
    return {"workload": f"{count} VR 2-skeleta, N~U{{20..80}}, seed 0, default options, queued, "
                        "Phase II launches timed",
            "columns": M, "batched_ms": statistics.median(ts) * 1e3, "reps": reps,
            "roofline": phase2_roofline(stats, M, f"k_phase2_group<{st.lanes}, {st.group_lds}>"
                                                  " + k_phase2_mid"),
            "num_iters": st.num_iters, "additions": st.additions,
            "merge_elems": st.merge_elems, "handoffs": st.handoffs}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--legs", default="all",
                    help="comma-separated subset of " + ",".join(LEGS) + " (default all); "
                         "'torus' is the headline and 'cpu' its CPU baseline")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--no-served", action="store_true")
    ap.add_argument("--no-builders", action="store_true",
                    help="skip the device-built filtration comparison")
    ap.add_argument("--no-twist", action="store_true",
                    help="skip the twist-on-the-same-GPU comparison")
    ap.add_argument("--no-tune", action="store_true",
                    help="batched leg with the default setting instead of the autotuner's")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="also run one 2 GiB device copy, the reference kernel that scales "
                         "rocprofv3 FETCH_SIZE / WRITE_SIZE readings to bytes")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; --gpus %d is not "
                                   "supported (independent matrices would shard across "
                                   "GPUs with no communication)" % args.gpus}))
        sys.exit(2)
    legs = set(DEFAULT_LEGS) if args.legs == "all" else set(args.legs.split(","))
    unknown = legs - set(LEGS)
    if unknown:
        ap.error("unknown legs: " + ",".join(sorted(unknown)))
    for flag, leg in (("no_cpu", "cpu"), ("no_batched", "batched"), ("no_served", "served"),
                      ("no_builders", "builders"), ("no_twist", "twist")):
        if getattr(args, flag):
            legs.discard(leg)
            if leg == "cpu":
                legs.discard("cubical_cpu")

    import torch

    import openph_amd
    from openph_amd import complexes as C

    openph_amd.lib()
    dev = torch.device("cuda:0")
    b = C.torus100()
    m, nnz = b.m, b.nnz
    gold = np.load(os.path.join(ROOT, "tests", "golden", "torus100.npz"))
    out = {}
    if "torus" in legs:
        wall, stats, tstats, ok = device_leg(b, "torus100", args.steps, args.warmup,
                                             (gold["low"], gold["ess"]))
        st = tstats[-1]
        dev_ms = statistics.median([s.device_ms for s in stats])
        kname = (f"k_phase2_group<{st.lanes}, {st.group_lds}>" if st.lanes > 1
                 else "k_phase2_lane")
        roof = phase2_roofline(tstats, m, kname)
        roof["timing"] = (f"HIP events around each Phase II launch: {len(tstats)} queued "
                          "reductions after the timed steps")
        # HBM bytes per launch of the same kernel from the committed rocprofv3
        # counter runs (FETCH_SIZE, WRITE_SIZE scaled by a plain copy)
        tfile = os.path.join(ROOT, "profiles", "pms_torus100_traffic.json")
        if os.path.exists(tfile):
            tj = json.load(open(tfile))
            if kname in tj.get("kernel", ""):
                roof["traffic"] = tj["traffic_bytes_per_launch"]
        value = m * args.steps / wall
        out.update({
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
            "roofline": roof,
            "hbm_alg": {"bytes": b_alg(st, m, nnz), "device_ms": dev_ms,
                        "GBps": b_alg(st, m, nnz) / (dev_ms * 1e-3) / 1e9 if dev_ms > 0 else 0.0},
            "pms": {"num_iters": st.num_iters, "additions": st.additions,
                    "schedule": stats[-1].schedule, "host_reads": stats[-1].host_reads,
                    "merge_elems": st.merge_elems, "compactions": st.compactions,
                    "pool_peak": st.pool_peak, "additions_per_s": st.additions * args.steps / wall,
                    "parity_vs_golden": ok},
        })

    if args.calibrate_copy:
        n = 512 << 20  # 2 GiB of int32, well beyond the 256 MB last-level cache
        src = torch.ones(n, dtype=torch.int32, device=dev)
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

    if "vr40" in legs:
        out["vr40"] = vr40_leg(cpu="cpu" in legs)
    if "dropin" in legs:
        print("bench: drop-in with trackers", file=sys.stderr, flush=True)
        out["dropin"] = dropin_leg(b)
    if "cubical" in legs:
        out["cubical64"] = cubical_leg(cpu="cpu" in legs)
    if "twist" in legs:
        print("bench: twist on the same GPU", file=sys.stderr, flush=True)
        out["gpu_twist"] = gpu_twist_leg(b)
    if "builders" in legs:
        print("bench: builders", file=sys.stderr, flush=True)
        out["builders"] = builder_leg()
    if "batched" in legs:
        print("bench: batched leg", file=sys.stderr, flush=True)
        out["batched"] = batched_leg(tune=not args.no_tune)
    if "batched_only" in legs:
        print("bench: batched call alone", file=sys.stderr, flush=True)
        out["batched_only"] = batched_only_leg()
    if "iterations" in legs:
        print("bench: per-iteration times", file=sys.stderr, flush=True)
        out["iterations"] = iterations_leg()
    if "phase_boundaries" in legs:
        print("bench: phase boundaries", file=sys.stderr, flush=True)
        out["phase_boundaries"] = phase_boundaries_leg()
    if "served" in legs:
        print("bench: served leg", file=sys.stderr, flush=True)
        out["served"] = served_leg()
    if "cpu" in legs and "torus" in legs:
        print("bench: cpu baseline", file=sys.stderr, flush=True)
        cb, cpu_low = cpu_baseline(b, "Torus-100")
        cb["low_sha_match"] = bool(np.array_equal(cpu_low, gold["low"]))
        out["cpu_baseline"] = cb
        out["speedup_vs_cpu"] = out["value"] / cb["value"]
    cub = out.get("cubical64")
    if cub and "_cpu_low" in cub:
        # the CPU twist's low against the GPU pms result of the same matrix
        import openph_amd as O
        low_g, _, _ = O.reduce(cub["_b"], "pms")
        cub["cpu_baseline"]["low_sha_match"] = bool(np.array_equal(cub.pop("_cpu_low"), low_g))
        cub.pop("_b")
        cub.pop("cpu_equals_gpu", None)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
