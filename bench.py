"""Benchmark: useful GB/s of the plain single-array APSEP solve on MI355X.

Workload (the headline configuration): N = 2^25 int32 keys in three classes
-- uniform random in [0, 2^31), strictly descending (a[i] = N - i), strictly
ascending (a[i] = i).  One step = one solve of each of the three arrays, all
resident in HBM before timing starts.  Useful bytes per solve = 2 * N * 4
(read the keys once, write the answers once).

Printed JSON line (rank 0, one line):
  value        aggregate useful GB/s over the three solves of a step
  roofline     the dominant kernel (the tile pass): achieved = 8 N / (its
               mean device time: after the timed steps, each kernel of the
               solve is launched `steps` times back to back per class with a
               HIP event pair around each batch on the same stream), peak =
               8000 GB/s HBM;
               traffic = its HBM bytes per launch from rocprofv3
               FETCH_SIZE/WRITE_SIZE runs scaled by a plain copy's counts
               (profiles/traffic.json; null when not measured)
  cpu_baseline the oracle's monotone stack (1 thread) on a bounded sample of
               the same inputs, timed on this host
Timing: HIP events on the stream the kernels are launched on (torch's current
stream, passed to the C ABI).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

PEAK_HBM_GBS = 8000.0
CLASSES = ("random", "descending", "ascending")


def make_inputs(torch, n: int, seed: int = 2024):
    """Synthetic inputs generated on the device (no host round trip)."""
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    idx = torch.arange(n, dtype=torch.int64, device="cuda")
    return {
        "random": torch.randint(0, 2**31 - 1, (n,), generator=g, dtype=torch.int32,
                                device="cuda"),
        "descending": (n - idx).to(torch.int32),
        "ascending": idx.to(torch.int32),
    }


def time_solves(torch, scratch, xs, outs, steps, warmup, stream):
    """Mean per-launch time (ms) of each class, HIP events on `stream`."""
    names = list(xs)
    for _ in range(warmup):
        for k in names:
            scratch.run(xs[k], outs[k], stream=stream)
    torch.cuda.synchronize()
    per = {k: [] for k in names}
    t_all0 = torch.cuda.Event(enable_timing=True)
    t_all1 = torch.cuda.Event(enable_timing=True)
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
          for _ in range(steps * len(names))]
    torch.cuda.synchronize()
    t_all0.record(stream)
    j = 0
    for _ in range(steps):
        for k in names:
            a, b = ev[j]
            a.record(stream)
            scratch.run(xs[k], outs[k], stream=stream)
            b.record(stream)
            j += 1
    t_all1.record(stream)
    torch.cuda.synchronize()
    scratch.check()
    # per-kernel device time: each kernel of the solve launched `steps` times
    # back to back per class, one HIP event pair per batch on the same stream
    kt = {k: scratch.time_kernels(xs[k], outs[k], steps, stream=stream) for k in names}
    scratch.check()
    j = 0
    for _ in range(steps):
        for k in names:
            a, b = ev[j]
            per[k].append(a.elapsed_time(b))
            j += 1
    return per, t_all0.elapsed_time(t_all1), kt


def copy_gbs(torch, n, stream, reps=20):
    src = torch.empty(n, dtype=torch.int32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    a = torch.cuda.Event(enable_timing=True)
    b = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    a.record(stream)
    for _ in range(reps):
        dst.copy_(src)
    b.record(stream)
    torch.cuda.synchronize()
    return 8.0 * n * reps / (a.elapsed_time(b) * 1e-3) / 1e9


def cpu_baseline(torch, xs, outs, budget_s: float):
    """The oracle's monotone stack, 1 thread, on the same inputs; also checks
    the GPU answers of the sample against it."""
    import numpy as np
    import oracle
    host = {k: v.cpu().numpy() for k, v in xs.items()}
    got = {k: v.cpu().numpy() for k, v in outs.items()}
    elems, secs, reps = 0, 0.0, 0
    ok = True
    while secs < budget_s or reps == 0:
        for k in CLASSES:
            t0 = time.perf_counter()
            exp = oracle.apsep(host[k])
            secs += time.perf_counter() - t0
            elems += host[k].shape[0]
            if reps == 0:
                ok = ok and bool(np.array_equal(exp, got[k]))
        reps += 1
    n = host[CLASSES[0]].shape[0]
    return {
        "value": round(8.0 * elems / secs / 1e9, 4),
        "unit": "GB/s",
        "cores": 1,
        "kind": "port",
        "sample": "%d passes over the 3 classes x N=%d (%.1f s of 1-thread "
                  "monotone stack, oracle/apsep_oracle.c -O3)" % (reps, n, secs),
        "ns_per_elem": round(secs / elems * 1e9, 3),
        "host_cores_available": len(os.sched_getaffinity(0)),
    }, ok


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=1 << 25)
    ap.add_argument("--algo", default="auto")
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--traffic", default=os.path.join(ROOT, "profiles", "traffic.json"),
                    help="rocprofv3 FETCH/WRITE summary (bytes per launch)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build measures one GPU only; --gpus %d is not "
              "supported (the solver does not shard across GPUs)" % args.gpus,
              file=sys.stderr)
        sys.exit(2)

    import torch
    import apsep_amd

    n = args.n
    stream = torch.cuda.current_stream()
    xs = make_inputs(torch, n)
    outs = {k: torch.empty(n, dtype=torch.int32, device="cuda") for k in CLASSES}
    torch.cuda.synchronize()
    with apsep_amd.Scratch(args.algo, n) as s:
        per, total_ms, kt = time_solves(torch, s, xs, outs, args.steps, args.warmup, stream)
        algo_name = {v: k for k, v in apsep_amd.ALGOS.items()}[s.algo]
    mean = {k: sum(v) / len(v) for k, v in per.items()}
    useful = 8.0 * n
    per_class = {k: round(useful / (mean[k] * 1e-3) / 1e9, 1) for k in CLASSES}
    step_ms = total_ms / args.steps
    value = 3 * useful / (step_ms * 1e-3) / 1e9
    runs = sum(t.runs for t in kt.values())
    tile_ms = sum(t.tile_ms for t in kt.values()) / runs
    fix_ms = sum(t.fix_ms for t in kt.values()) / runs
    levels_ms = sum(t.levels_ms for t in kt.values()) / runs
    achieved = useful / (tile_ms * 1e-3) / 1e9

    traffic = None
    if os.path.exists(args.traffic):
        try:
            with open(args.traffic) as f:
                t = json.load(f)
            if t.get("n") == n and t.get("algo") == algo_name:
                traffic = t.get("tile_pass_bytes_per_launch")
        except Exception:
            traffic = None

    rec = {
        "metric": "useful GB/s at N=32M int32 (% of HBM peak)",
        "value": round(value, 1),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(step_ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32",
        "data": "synthetic (device-generated: uniform random seed 2024, "
                "strictly descending, strictly ascending)",
        "config": {"workload": "plain single-array APSEP, N=2^25 int32, "
                               "3 classes per step (random/descending/ascending)",
                   "n": n, "algo": algo_name},
        "pct_of_hbm_peak": round(100.0 * value / PEAK_HBM_GBS, 2),
        "per_class_gbs": per_class,
        "per_class_us": {k: round(mean[k] * 1e3, 2) for k in CLASSES},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1),
                     "peak": PEAK_HBM_GBS, "unit": "GB/s",
                     "frac": round(achieved / PEAK_HBM_GBS, 4), "traffic": traffic,
                     "kernel": "k_tile_pass (mean over the 3 classes)",
                     "algorithmic_bytes_per_launch": int(useful)},
        "kernel_us": {"tile_pass": round(tile_ms * 1e3, 2),
                      "level_build": round(levels_ms * 1e3, 2),
                      "fix_pass": round(fix_ms * 1e3, 2), "launches": int(runs),
                      "tile_pass_per_class": {k: round(t.tile_ms / t.runs * 1e3, 2)
                                              for k, t in kt.items()}},
        "copy_gbs": round(copy_gbs(torch, n, stream), 1),
    }
    if not args.no_cpu:
        cb, ok = cpu_baseline(torch, xs, outs, args.cpu_seconds)
        rec["cpu_baseline"] = cb
        rec["parity_checked"] = ok
        if not ok:
            print("bench.py: GPU answers differ from the oracle", file=sys.stderr)
    print(json.dumps(rec), flush=True)


if __name__ == "__main__":
    main()
