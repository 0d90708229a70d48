"""Benchmark: useful GB/s of the plain single-array APSEP solve on MI355X,
plus the batched and served forms.

Headline workload: N = 2^25 int32 keys in three classes -- uniform random
in [0, 2^31), strictly descending (a[i] = N - i), strictly ascending
(a[i] = i).  One step = one solve of each of the three arrays (each with its
own input and output buffer), all resident in HBM before timing starts.
Useful bytes per solve = 2 * N * 4 (read the keys once, write the answers
once).

Printed JSON line (rank 0, one line):
  value        aggregate useful GB/s over the three solves of a step
  settle_s     untimed solves run before the warmup steps so that the GPU
               reaches its sustained clock (the first milliseconds after an
               idle GPU run slower)
  roofline     the dominant kernel (the tile pass): achieved = 8 N / its
               mean device time, measured as in the timed steps' cache
               regime -- each kernel of the solve launched back to back over
               the three classes in turn, each class with its own scratch,
               input and output (768 MB, three times the last-level cache),
               one HIP event pair around each kernel's batch on the stream
               the kernels run on; peak = 8000 GB/s HBM; traffic = its HBM
               bytes per launch from rocprofv3 FETCH_SIZE/WRITE_SIZE runs
               scaled by a plain copy's counts (profiles/traffic.json; null
               when not measured).  warm_* beside it: the same kernels
               re-run on one array (its data partly cache-resident)
  cpu_baseline the oracle's monotone stack (1 thread) on a bounded sample of
               the same inputs, timed on this host
  per_class_*  each class alone: solves back to back between one event pair,
               over three input/output pairs of the class in turn (_warm: on
               one pair)
  structured   the tie- and structure-heavy arrays at N = 2^28 (uniform
               [0,16), sawtooth i mod 4096, block-descending runs of 2^16):
               useful GB/s per array, the solve's kernels, oracle parity
  batch        the batched configuration (8192 segments, log-uniform lengths
               64..2^20, ~8.7e8 keys): arrays/s and useful GB/s of one
               segmented solve, against one apsep_run and one apsep_launch
               per array on the same arrays
  serve        the served configuration (Poisson arrivals, lengths
               log-uniform 2^10..2^24, classes 60/20/20, 64 GB pool) at a few
               offered rates: jobs/s and p50/p99 latency, against one
               apsep_run and one apsep_launch per job on the same job mix
Timing: HIP events on the stream the kernels are launched on (torch's current
stream, passed to the C ABI); the timed region holds the `steps` steps and no
other event.
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


def settle(torch, scratch, xs, outs, stream, seconds):
    """Untimed solves for `seconds` of wall time (at least one round)."""
    names = list(xs)
    t0 = time.perf_counter()
    rounds = 0
    while rounds == 0 or time.perf_counter() - t0 < seconds:
        for _ in range(10):
            for k in names:
                scratch.run(xs[k], outs[k], stream=stream)
        torch.cuda.synchronize()
        rounds += 10
    return rounds


def time_solves(torch, apsep_amd, scratch, xs, outs, steps, warmup, stream, per_reps):
    """Timed region: `steps` steps of one solve per class, back to back on
    `stream` with nothing between the solves, one HIP event pair around the
    whole region.  Then, outside it: the per-class breakdown (`per_reps`
    solves of one class back to back between one event pair, over three
    distinct input/output pairs of that class in turn, and again on one pair
    (warm)), each kernel of the solve launched over the three classes in turn
    (one scratch per class), and the same kernels re-run on one class at a
    time (warm)."""
    names = list(xs)
    for _ in range(warmup):
        for k in names:
            scratch.run(xs[k], outs[k], stream=stream)
    torch.cuda.synchronize()
    t_all0 = torch.cuda.Event(enable_timing=True)
    t_all1 = torch.cuda.Event(enable_timing=True)
    t_all0.record(stream)
    for _ in range(steps):
        for k in names:
            scratch.run(xs[k], outs[k], stream=stream)
    t_all1.record(stream)
    torch.cuda.synchronize()
    scratch.check()
    per, per_warm = {}, {}
    for k in names:
        ins = [xs[k]] + [xs[k].clone() for _ in range(2)]
        res = [outs[k]] + [torch.empty_like(outs[k]) for _ in range(2)]
        torch.cuda.synchronize()
        for m, dst in ((3, per), (1, per_warm)):
            a = torch.cuda.Event(enable_timing=True)
            b = torch.cuda.Event(enable_timing=True)
            a.record(stream)
            for r in range(per_reps):
                scratch.run(ins[r % m], res[r % m], stream=stream)
            b.record(stream)
            torch.cuda.synchronize()
            dst[k] = a.elapsed_time(b) / per_reps
        del ins, res
    scratch.check()
    n = xs[names[0]].numel()
    algo = {v: k for k, v in apsep_amd.ALGOS.items()}[scratch.algo]
    rot_s = [apsep_amd.Scratch(algo, n) for _ in names]
    try:
        reps = per_reps * len(names)
        rot = apsep_amd.time_kernels_rotating(rot_s, [xs[k] for k in names],
                                              [outs[k] for k in names], reps, stream=stream)
        for r in rot_s:
            r.check()
    finally:
        for r in rot_s:
            r.close()
    warm = {k: scratch.time_kernels(xs[k], outs[k], per_reps, stream=stream) for k in names}
    scratch.check()
    return per, per_warm, t_all0.elapsed_time(t_all1), rot, warm


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


def batch_inputs(torch, nseg: int, lo: int, hi: int, seed: int):
    """Log-uniform segment lengths floor(exp(U(ln lo, ln hi))), class per
    segment uniform over random / descending / ascending; keys built on the
    device."""
    import numpy as np
    rng = np.random.default_rng(seed)
    lengths = np.floor(np.exp(rng.uniform(np.log(lo), np.log(hi), size=nseg))).astype(np.int64)
    cls = rng.integers(0, 3, size=nseg)
    off = np.zeros(nseg + 1, dtype=np.int64)
    np.cumsum(lengths, out=off[1:])
    total = int(off[-1])
    dev = "cuda"
    L = torch.from_numpy(lengths).to(dev)
    seg = torch.repeat_interleave(torch.arange(nseg, device=dev), L)
    local = torch.arange(total, device=dev, dtype=torch.int64) - torch.from_numpy(off[:-1]).to(dev)[seg]
    c = torch.from_numpy(cls).to(dev)[seg]
    g = torch.Generator(device=dev)
    g.manual_seed(seed)
    rnd = torch.randint(0, 2**31 - 1, (total,), generator=g, dtype=torch.int32, device=dev)
    desc = (L[seg] - local).to(torch.int32)
    x = torch.where(c == 0, rnd, torch.where(c == 1, desc, local.to(torch.int32)))
    del seg, local, c, rnd, desc
    return x.contiguous(), off, lengths, cls


def _timed(torch, stream, fn, reps):
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    ev0.record(stream)
    for _ in range(reps):
        fn()
    ev1.record(stream)
    torch.cuda.synchronize()
    return ev0.elapsed_time(ev1) / reps


def _packed(torch, x, off, lengths):
    """Every array of a CSR batch in its own 16-byte aligned slot (as a
    caller with one buffer per array has them); returns (keys, answers,
    slot starts)."""
    import numpy as np
    nseg = lengths.shape[0]
    pstart = np.zeros(nseg + 1, dtype=np.int64)
    np.cumsum((lengths + 3) // 4 * 4, out=pstart[1:])
    seg = torch.repeat_interleave(torch.arange(nseg, device="cuda"),
                                  torch.from_numpy(lengths).cuda())
    offs = torch.from_numpy(off).cuda()
    dst = torch.arange(x.numel(), device="cuda") - offs[:-1][seg] + \
        torch.from_numpy(pstart[:-1]).cuda()[seg]
    xp = torch.zeros(int(pstart[-1]), dtype=torch.int32, device="cuda")
    xp[dst] = x
    del seg, dst, offs
    return xp, torch.empty_like(xp), pstart


def per_array_baselines(torch, apsep_amd, stream, xp, outp, pstart, lengths, launch_count):
    """One solve per array: apsep_run on one persistent scratch, all arrays
    back to back on `stream` (ms for all), and apsep_launch (allocate,
    solve, synchronise, free: the reference's launchWSTL / launchSPT) for the
    first `launch_count` arrays (ms per array)."""
    nseg = lengths.shape[0]
    xb, ob = xp.data_ptr(), outp.data_ptr()
    with apsep_amd.Scratch("auto", int(lengths.max())) as s:
        sh = stream.cuda_stream

        def per_array():
            for k in range(nseg):
                s.run_ptr(xb + 4 * int(pstart[k]), ob + 4 * int(pstart[k]), int(lengths[k]), sh)
        per_array()
        run_ms = _timed(torch, stream, per_array, 1)
    m = min(launch_count, nseg)

    def per_launch():
        for k in range(m):
            apsep_amd.launch("auto", xb + 4 * int(pstart[k]), ob + 4 * int(pstart[k]),
                             int(lengths[k]))
    launch_ms = _timed(torch, stream, per_launch, 1) / m
    return run_ms, launch_ms


def run_batch(args, torch, apsep_amd, cpu: bool):
    """Segmented batch: 8192 segments with log-uniform lengths in [64, 2^20]
    (about 8.7e8 keys), one class per segment.  The batched solve (the
    autotuner's tile shape and launch sequence, eager and hipGraph replay;
    one persistent launch per batch timed beside it) against one solve per
    array on the same arrays (apsep_run on a persistent scratch, and
    apsep_launch)."""
    import numpy as np
    x, off, lengths, cls = batch_inputs(torch, args.nseg, 64, 1 << 20, 2024)
    total, nseg = int(off[-1]), args.nseg
    offs = torch.from_numpy(off).cuda()
    out = torch.empty(total, dtype=torch.int32, device="cuda")
    res = {}
    stream = torch.cuda.Stream()  # graphs need a capturable (non-null) stream
    with apsep_amd.BatchPlan(total, nseg) as p, torch.cuda.stream(stream):
        run = lambda: p.run(x, offs, out, total=total, stream=stream)
        tuned = p.tune(x, offs, out, total=total, stream=stream)
        res["tuned_mode"] = "graph" if tuned[1] < tuned[0] else "eager"
        res["tile_lanes"] = p.tile_lanes
        res["tuned_persistent"] = p.tuned_persistent(int(total - 1).bit_length())
        p.use_graph(False)
        # the same batch as one persistent launch (tuned tile shape)
        p.set_persistent(True)
        run()
        res["persistent_ms"] = _timed(torch, stream, run, max(3, args.batch_steps // 2))
        p.check()
        p.set_persistent(res["tuned_persistent"])
        for _ in range(args.batch_warmup):
            run()
        res["eager_ms"] = _timed(torch, stream, run, args.batch_steps)
        p.use_graph(True)
        run()
        res["graph_ms"] = _timed(torch, stream, run, args.batch_steps)
    torch.cuda.synchronize()
    got = out.cpu().numpy() if cpu else None
    xp, outp, pstart = _packed(torch, x, off, lengths)
    run_ms, launch_ms = per_array_baselines(torch, apsep_amd, stream, xp, outp, pstart,
                                            lengths, nseg if args.workload == "batch" else 1024)
    best = min(res["eager_ms"], res["graph_ms"])
    useful = 8.0 * total
    rec = {
        "workload": "segmented CSR batch, %d segments, log-uniform lengths 64..2^20, "
                    "%d keys" % (nseg, total),
        "arrays_per_s": round(nseg / (best * 1e-3), 1),
        "useful_GBps": round(useful / (best * 1e-3) / 1e9, 1),
        "ms_per_batch": round(best, 4),
        "eager_ms": round(res["eager_ms"], 4), "graph_ms": round(res["graph_ms"], 4),
        "autotuned_mode": res["tuned_mode"], "autotuned_tile_lanes": res["tile_lanes"],
        "autotuned_persistent": bool(res["tuned_persistent"]),
        "persistent_launch_ms": round(res["persistent_ms"], 4),
        "per_array_run_ms": round(run_ms, 3),
        "per_array_launch_ms_est": round(launch_ms * nseg, 2),
        "per_array_launch_sampled": min(nseg, nseg if args.workload == "batch" else 1024),
        "speedup_vs_per_array_run": round(run_ms / best, 2),
        "speedup_vs_per_array_launch": round(launch_ms * nseg / best, 2),
    }
    if cpu:
        import oracle
        host = x.cpu().numpy()
        threads = len(os.sched_getaffinity(0))
        t0 = time.perf_counter()
        exp = oracle.apsep_segmented(host, off, False, threads=threads)
        secs = time.perf_counter() - t0
        rec["cpu_baseline"] = {
            "value": round(useful / secs / 1e9, 3), "unit": "GB/s", "cores": threads,
            "kind": "port",
            "sample": "the whole batch once, segments as jobs on a %d-thread pool "
                      "(oracle/apsep_oracle.c -O3), %.2f s" % (threads, secs)}
        rec["parity_checked"] = bool(np.array_equal(exp, got))
        outp_h = outp.cpu().numpy()
        rec["per_array_parity_checked"] = bool(all(
            np.array_equal(outp_h[pstart[k]:pstart[k] + lengths[k]], got[off[k]:off[k + 1]])
            for k in range(0, nseg, 97)))
    del x, out, xp, outp
    torch.cuda.empty_cache()
    return rec


def bench_batch(args, torch, apsep_amd):
    rec = run_batch(args, torch, apsep_amd, not args.no_cpu)
    line = {
        "metric": "batched useful GB/s (8192 segments, log-uniform lengths 64..2^20)",
        "value": rec["useful_GBps"], "unit": "GB/s", "n_gpus": 1,
        "steps": args.batch_steps, "warmup": args.batch_warmup,
        "ms_per_step": rec["ms_per_batch"], "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "int32",
        "data": "synthetic (numpy seed 2024 lengths/classes, device-generated keys)",
        "config": {"workload": rec["workload"], "segments": args.nseg},
    }
    line.update({k: v for k, v in rec.items() if k != "workload"})
    print(json.dumps(line), flush=True)


def serve_payloads(torch, P, seed):
    """P distinct jobs: lengths log-uniform in [2^10, 2^24], classes 60/20/20
    random / descending / ascending, keys generated on the device into one
    buffer (16-byte aligned slots)."""
    import numpy as np
    rng = np.random.default_rng(seed)
    lengths = np.floor(np.exp(rng.uniform(np.log(1 << 10), np.log(1 << 24), size=P))).astype(np.int64)
    cls = rng.choice(3, size=P, p=[0.6, 0.2, 0.2])
    off = np.zeros(P + 1, dtype=np.int64)
    np.cumsum((lengths + 3) // 4 * 4, out=off[1:])
    total = int(off[-1])
    dev = "cuda"
    Lp = torch.from_numpy((lengths + 3) // 4 * 4).to(dev)
    seg = torch.repeat_interleave(torch.arange(P, device=dev), Lp)
    local = torch.arange(total, device=dev, dtype=torch.int64) - torch.from_numpy(off[:-1]).to(dev)[seg]
    L = torch.from_numpy(lengths).to(dev)
    c = torch.from_numpy(cls).to(dev)[seg]
    g = torch.Generator(device=dev)
    g.manual_seed(seed)
    rnd = torch.randint(0, 2**31 - 1, (total,), generator=g, dtype=torch.int32, device=dev)
    src = torch.where(c == 0, rnd, torch.where(c == 1, (L[seg] - local).to(torch.int32),
                                               local.to(torch.int32))).contiguous()
    del seg, local, c, rnd
    torch.cuda.synchronize()
    return src, off, lengths, cls


def run_serve(args, torch, apsep_amd, cpu: bool):
    """Served stream of independent jobs: Poisson arrivals at swept offered
    rates, lengths log-uniform in [2^10, 2^24], classes 60/20/20 random /
    descending / ascending, a 64 GB device pool.  Jobs' keys are resident in
    device memory when they arrive (a fixed set of distinct payloads, reused
    round-robin); latency runs from the scheduled arrival (so a job refused
    by admission waits in the client backlog) to the moment the server
    observes its batch finished.  Baselines on the same job sequence: one
    apsep_run per job back to back, and apsep_launch per job."""
    import numpy as np
    rng = np.random.default_rng(7)
    P = args.payloads
    src, off, lengths, cls = serve_payloads(torch, P, 7)
    base = src.data_ptr()
    rates = [float(r) for r in args.rates.split(",")]
    J = args.jobs
    sweep = []
    sample_answers = {}
    ptr_arr = np.array([base + 4 * int(off[k]) for k in range(P)], dtype=np.uint64)
    len_arr = lengths.astype(np.int64)
    with apsep_amd.Server(64 << 30, 1 << 28, 4096) as srv:
        for rate in rates:
            arr = (time.monotonic_ns() / 1e3 + 2e5
                   + np.cumsum(rng.exponential(1e6 / rate, size=J)))
            which = np.arange(J) % P
            jp, jl = ptr_arr[which], len_arr[which]
            nxt = rel = 0      # jobs submitted / released
            id0 = None         # id of this rate's first job (ids are consecutive)
            lat = np.zeros(J)
            refused = 0
            t_last = arr[0]
            while rel < J:
                now = time.monotonic_ns() / 1e3
                m = int(np.searchsorted(arr, now, side="right"))
                if m > nxt:
                    first, acc = srv.submit_many(jp[nxt:m], jl[nxt:m])
                    if id0 is None:
                        id0 = first
                    refused += acc < m - nxt
                    nxt += acc
                srv.poll()
                if id0 is not None:
                    d = srv.done_upto() - id0
                    if d > rel:
                        _, done = srv.job_times(id0 + rel, d - rel)
                        lat[rel:d] = done - arr[rel:d]
                        t_last = float(done[-1])
                        srv.release_upto(id0 + d)
                        rel = d
                if nxt < J and rel == nxt:
                    wait = (arr[nxt] - time.monotonic_ns() / 1e3) * 1e-6
                    if wait > 2e-4:
                        time.sleep(min(wait - 1e-4, 1e-3))
            span_s = (t_last - arr[0]) * 1e-6
            done_elems = int(jl.sum())
            lat_ms = lat * 1e-3
            sweep.append({"offered_jobs_per_s": rate,
                          "jobs_per_s": round(J / span_s, 1),
                          "keys_per_s": round(done_elems / span_s, 1),
                          "useful_GBps": round(8.0 * done_elems / span_s / 1e9, 1),
                          "p50_ms": round(float(np.percentile(lat_ms, 50)), 3),
                          "p99_ms": round(float(np.percentile(lat_ms, 99)), 3),
                          "admission_refusals": int(refused)})
            print("serve rate %.0f/s: %s" % (rate, json.dumps(sweep[-1])), file=sys.stderr,
                  flush=True)
        if cpu:
            # parity, outside the timed sweeps: every distinct payload once
            # through the same server, all answers kept and compared below
            k0 = 0
            while k0 < P:  # (in pieces when the pool cannot hold all at once)
                first, acc = srv.submit_many(ptr_arr[k0:], len_arr[k0:])
                if acc == 0:
                    raise RuntimeError("bench.py: serve pool smaller than one payload")
                srv.drain()
                for j in range(acc):
                    sample_answers[k0 + j] = srv.answers(first + j)
                srv.release_upto(first + acc)
                k0 += acc
    # the same job sequence, one plain solve per job
    which = np.arange(J) % P
    stream = torch.cuda.current_stream()
    outp = torch.empty_like(src)
    xb, ob = src.data_ptr(), outp.data_ptr()
    # the distinct payloads as back-to-back job batches of the server's
    # size limit (no arrivals, no admission): what the kernels reach on this
    # mix, the ceiling for the served rate
    mix_lanes = 256
    batches, cur, curn = [], [], 0
    for k in range(P):
        if cur and curn + int(lengths[k]) > (1 << 28):
            batches.append(cur)
            cur, curn = [], 0
        cur.append(k)
        curn += int(lengths[k])
    batches.append(cur)
    xs = [src[off[k]:off[k] + lengths[k]] for k in range(P)]
    os_ = [outp[off[k]:off[k] + lengths[k]] for k in range(P)]
    with apsep_amd.BatchPlan((1 << 28) + 4096 * 8192, 4096) as bp:
        bp.set_tile_lanes(mix_lanes)
        tabs = [bp.job_table([xs[k] for k in b], [os_[k] for k in b]) for b in batches]

        def mix_pass():
            for (tb, nt), b in zip(tabs, batches):
                bp.run_jobs(None, None, stream=stream, table=tb, tiles=nt)
        for _ in range(3):
            mix_pass()
        mix_ms = _timed(torch, stream, mix_pass, 5)
    mix_keys = int(lengths.sum())
    del xs, os_, tabs
    with apsep_amd.Scratch("auto", int(lengths.max())) as s:
        sh = stream.cuda_stream

        def per_job():
            for j in range(J):
                k = which[j]
                s.run_ptr(xb + 4 * int(off[k]), ob + 4 * int(off[k]), int(lengths[k]), sh)
        per_job()
        run_ms = _timed(torch, stream, per_job, 1)
    m = min(J, 500)

    def per_launch():
        for j in range(m):
            k = which[j]
            apsep_amd.launch("auto", xb + 4 * int(off[k]), ob + 4 * int(off[k]),
                             int(lengths[k]))
    launch_ms = _timed(torch, stream, per_launch, 1) / m
    best = max(sweep, key=lambda r: r["jobs_per_s"])
    run_jps = J / (run_ms * 1e-3)
    launch_jps = 1.0 / (launch_ms * 1e-3)
    rec = {
        "workload": "served stream, %d jobs per rate, %d distinct payloads, mean %.0f keys, "
                    "64 GB pool" % (J, P, float(lengths[which].mean())),
        "best_jobs_per_s": best["jobs_per_s"],
        "best_useful_GBps": best["useful_GBps"],
        "sweep": sweep,
        "batched_same_mix": {
            "jobs_per_s": round(P / (mix_ms * 1e-3), 1),
            "useful_GBps": round(8.0 * mix_keys / (mix_ms * 1e-3) / 1e9, 1),
            "sample": "the %d distinct payloads as %d back-to-back job batches of <= 2^28 keys, "
                      "%d-lane tiles" % (P, len(batches), mix_lanes)},
        "served_vs_batched_same_mix": round(best["useful_GBps"] /
                                            (8.0 * mix_keys / (mix_ms * 1e-3) / 1e9), 3),
        "per_job_run_jobs_per_s": round(run_jps, 1),
        "per_job_launch_jobs_per_s": round(launch_jps, 1),
        "per_job_launch_sampled": m,
        "speedup_vs_per_job_run": round(best["jobs_per_s"] / run_jps, 2),
        "speedup_vs_per_job_launch": round(best["jobs_per_s"] / launch_jps, 2),
    }
    if cpu:
        import oracle
        host = src.cpu().numpy()
        hl = np.zeros(P + 1, dtype=np.int64)
        np.cumsum(lengths, out=hl[1:])
        keys = np.concatenate([host[off[k]:off[k] + lengths[k]] for k in range(P)])
        threads = len(os.sched_getaffinity(0))
        t0 = time.perf_counter()
        exp = oracle.apsep_segmented(keys, hl, False, threads=threads)
        secs = time.perf_counter() - t0
        rec["cpu_baseline"] = {
            "value": round(P / secs, 1), "unit": "jobs/s", "cores": threads, "kind": "port",
            "sample": "the %d distinct payloads once, as independent jobs on a %d-thread "
                      "pool (oracle/apsep_oracle.c -O3), %.2f s" % (P, threads, secs)}
        rec["parity_checked"] = bool(sample_answers) and bool(all(
            np.array_equal(v.cpu().numpy(), exp[hl[k]:hl[k + 1]])
            for k, v in sample_answers.items()))
        rec["parity_jobs_checked"] = len(sample_answers)  # (every distinct payload)
    del src, outp
    torch.cuda.empty_cache()
    return rec


def bench_serve(args, torch, apsep_amd):
    rec = run_serve(args, torch, apsep_amd, not args.no_cpu)
    line = {
        "metric": "served jobs/s (Poisson arrivals, lengths log-uniform 2^10..2^24, 64 GB pool)",
        "value": rec["best_jobs_per_s"], "unit": "jobs/s", "n_gpus": 1, "steps": args.jobs,
        "warmup": 0, "ms_per_step": None, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "int32",
        "data": "synthetic (numpy seed 7 lengths/classes/arrivals, device-generated keys, "
                "%d distinct payloads reused round-robin)" % args.payloads,
        "config": {"workload": rec["workload"], "rates": args.rates},
    }
    line.update({k: v for k, v in rec.items() if k != "workload"})
    print(json.dumps(line), flush=True)


def bench_tune(args, torch, apsep_amd):
    """The autotuner's choices per size class: for batches of about 2^c keys
    (c = 14 .. 28; log-uniform segment lengths 64 .. min(2^20, 2^c / 4),
    class per segment uniform over random / descending / ascending), every
    tile shape x launch sequence (kernel sequence, or one persistent launch:
    "p") x launch mode (e = direct, g = graph replay) timed (median of
    `steps` runs), and what apsep_batch_plan_tune keeps for the class."""
    import numpy as np
    rows = []
    stream = torch.cuda.Stream()
    for c in range(14, 29):
        target = 1 << c
        hi = max(64, min(1 << 20, target >> 2))
        mean = (hi - 64) / np.log(hi / 64.0) if hi > 64 else 64
        nseg = max(1, min(200000, int(round(target / mean))))
        x, off, lengths, _ = batch_inputs(torch, nseg, 64, hi, 1000 + c)
        total = int(off[-1])
        offs = torch.from_numpy(off).cuda()
        out = torch.empty(total, dtype=torch.int32, device="cuda")
        row = {"target_log2": c, "class": int(total - 1).bit_length(), "keys": total,
               "segments": nseg, "times_ms": {}}
        with apsep_amd.BatchPlan(total, nseg) as p, torch.cuda.stream(stream):
            for lanes in (128, 256, 512):
                for persist in (0, 1):
                    for graph in (0, 1):
                        p.set_tile_lanes(lanes)
                        p.set_persistent(bool(persist))
                        p.use_graph(bool(graph))
                        run = lambda: p.run(x, offs, out, total=total, stream=stream)
                        run()
                        run()
                        ts = []
                        for _ in range(args.steps):
                            ts.append(_timed(torch, stream, run, 1))
                        row["times_ms"]["%d%s%s" % (lanes, "p" if persist else "",
                                                    "g" if graph else "e")] = \
                            round(sorted(ts)[len(ts) // 2], 4)
            p.check()
            p.tune(x, offs, out, total=total, stream=stream, reps=max(3, args.steps // 2))
            cls = int(total - 1).bit_length()  # the library's size class of `total`
            lanes, graph, ms = p.tuned(cls)
            row["tuned"] = {"lanes": lanes, "persistent": p.tuned_persistent(cls),
                            "graph": bool(graph), "ms": round(ms, 4)}
        best = min(row["times_ms"], key=row["times_ms"].get)
        row["best"] = best
        row["best_GBps"] = round(8.0 * total / (row["times_ms"][best] * 1e-3) / 1e9, 1)
        rows.append(row)
        print("tune class %d: %s" % (c, json.dumps(row)), file=sys.stderr, flush=True)
        del x, out, offs
        torch.cuda.empty_cache()
    print(json.dumps({
        "metric": "segmented batch time per size class, every tile shape x launch sequence x mode",
        "value": None, "unit": "ms", "n_gpus": 1, "steps": args.steps, "warmup": 2,
        "ms_per_step": None, "higher_is_better": False, "scaling": "weak",
        "vs_baseline": None, "dtype": "int32",
        "data": "synthetic (numpy lengths/classes, device-generated keys)",
        "config": {"workload": "autotuner size classes 2^14..2^28"}, "rows": rows}), flush=True)


def structured_inputs(torch, n: int, seed: int = 11):
    """The tie- and structure-heavy single arrays, built on the device:
    uniform keys in [0, 16), a sawtooth i mod 4096, and block-descending
    runs of 2^16 keys (a[i] = off[r] + 2^16 - 1 - i mod 2^16, off[r] uniform
    in [0, 2^31 - 2^16))."""
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    i = torch.arange(n, dtype=torch.int64, device="cuda")
    off = torch.randint(0, 2**31 - 2**16, ((n + 65535) >> 16,), generator=g,
                        dtype=torch.int64, device="cuda")
    xs = {
        "dup16": torch.randint(0, 16, (n,), generator=g, dtype=torch.int32, device="cuda"),
        "sawtooth": (i % 4096).to(torch.int32),
        "blockdesc": (off[i >> 16] + (65535 - (i & 65535))).to(torch.int32),
    }
    del i, off
    return xs


def run_structured(args, torch, apsep_amd, cpu: bool):
    """Solve time of each structured input at N = 2^28 (`structured_reps`
    solves back to back between one event pair, after two untimed ones),
    the split into the solve's kernels, and with `cpu` a check of every
    answer against the oracle, whose time on the same arrays is the CPU
    baseline."""
    n = args.structured_n
    stream = torch.cuda.current_stream()
    xs = structured_inputs(torch, n)
    out = torch.empty(n, dtype=torch.int32, device="cuda")
    rows = {}
    ok, secs = True, 0.0
    with apsep_amd.Scratch(args.algo, n) as s:
        for k, x in xs.items():
            for _ in range(2):
                s.run(x, out, stream=stream)
            ms = _timed(torch, stream, lambda: s.run(x, out, stream=stream),
                        args.structured_reps)
            t = s.time_kernels(x, out, args.structured_reps, stream=stream)
            s.check()
            rows[k] = {"us": round(ms * 1e3, 1),
                       "useful_GBps": round(8.0 * n / (ms * 1e-3) / 1e9, 1),
                       "kernel_us": _kernel_us(t)}
            print("structured %s: %s" % (k, json.dumps(rows[k])), file=sys.stderr, flush=True)
            if cpu:
                import numpy as np
                import oracle
                host = x.cpu().numpy()
                t0 = time.perf_counter()
                exp = oracle.apsep(host)
                secs += time.perf_counter() - t0
                ok = ok and bool(np.array_equal(exp, out.cpu().numpy()))
                del host, exp
    rec = {"workload": "tie- and structure-heavy single arrays, N=2^%d int32: uniform [0,16), "
                       "sawtooth i mod 4096, block-descending runs of 2^16"
                       % (n.bit_length() - 1),
           "n": n, "per_kind": rows,
           "useful_GBps": round(3 * 8.0 * n / (sum(r["us"] for r in rows.values()) * 1e-6)
                                / 1e9, 1)}
    if cpu:
        rec["parity_checked"] = ok
        rec["cpu_baseline"] = {"value": round(3 * 8.0 * n / secs / 1e9, 3), "unit": "GB/s",
                               "cores": 1, "kind": "port",
                               "sample": "each of the 3 arrays once (%.1f s of 1-thread "
                                         "monotone stack, oracle/apsep_oracle.c -O3)" % secs}
        if not ok:
            print("bench.py: structured answers differ from the oracle", file=sys.stderr)
    del xs, out
    torch.cuda.empty_cache()
    return rec


def bench_structured(args, torch, apsep_amd):
    rec = run_structured(args, torch, apsep_amd, not args.no_cpu)
    print(json.dumps({
        "metric": "useful GB/s, structure-heavy single arrays at N=2^28 int32",
        "value": rec["useful_GBps"], "unit": "GB/s", "n_gpus": 1,
        "steps": args.structured_reps, "warmup": 2, "ms_per_step": None,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "int32",
        "data": "synthetic (device-generated, torch seed 11)",
        "config": {"workload": rec["workload"], "n": rec["n"], "algo": args.algo},
        **{k: v for k, v in rec.items() if k not in ("workload", "n")}}), flush=True)


def bench_sweep(args, torch, apsep_amd):
    """Useful GB/s of the plain solve for N = 2^14 .. 2^27 on the three
    classes in three cache regimes, beside a plain device copy of the same
    N in the same regimes (median of --steps timings, only the solve or the
    copy inside the events):
      warm           repeated solves of the same array;
      flushed        a 1 GiB buffer rewritten before every solve, which
                     evicts the 256 MB last-level cache and leaves it full
                     of dirty lines to be written back during the solve;
      flushed_read   a 1 GiB buffer read (summed) before every solve, which
                     evicts the cache and leaves it clean."""
    stream = torch.cuda.current_stream()
    flush = torch.empty(1 << 28, dtype=torch.int32, device="cuda")
    flush.fill_(1)
    acc = torch.empty((), dtype=torch.int64, device="cuda")

    def timed(fn, regime):
        evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
               for _ in range(args.steps)]
        torch.cuda.synchronize()
        for i in range(args.steps):
            if regime == "flushed":
                flush.fill_(i)
            elif regime == "flushed_read":
                torch.sum(flush, dim=0, dtype=torch.int64, out=acc)
            evs[i][0].record(stream)
            fn()
            evs[i][1].record(stream)
        torch.cuda.synchronize()
        return sorted(a.elapsed_time(b) for a, b in evs)[args.steps // 2]

    regimes = ("warm", "flushed", "flushed_read")
    rows = []
    with apsep_amd.Scratch(args.algo, 1 << 27) as s:
        for lg in range(14, 28):
            n = 1 << lg
            xs = make_inputs(torch, n)
            out = torch.empty(n, dtype=torch.int32, device="cuda")
            row = {"n": n}
            for k in CLASSES:
                for _ in range(3):
                    s.run(xs[k], out, stream=stream)
                row[k] = {}
                for reg in regimes:
                    ms = timed(lambda: s.run(xs[k], out, stream=stream), reg)
                    row[k][reg + "_GBps"] = round(8.0 * n / (ms * 1e-3) / 1e9, 1)
                    row[k][reg + "_us"] = round(ms * 1e3, 2)
            s.check()
            src = xs["random"]
            row["copy"] = {}
            for reg in regimes:
                ms = timed(lambda: out.copy_(src), reg)
                row["copy"][reg + "_GBps"] = round(8.0 * n / (ms * 1e-3) / 1e9, 1)
            rows.append(row)
            print("sweep n=2^%d %s" % (lg, json.dumps(row)), file=sys.stderr, flush=True)
    best = max(r["random"]["warm_GBps"] for r in rows)
    print(json.dumps({
        "metric": "useful GB/s sweep, N=2^14..2^27 x 3 classes (median of %d solves)" % args.steps,
        "value": rows[-3]["random"]["warm_GBps"], "unit": "GB/s", "n_gpus": 1,
        "steps": args.steps, "warmup": 3, "ms_per_step": None, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "int32",
        "data": "synthetic (device-generated, as the headline)",
        "config": {"workload": "plain solve size sweep", "algo": args.algo},
        "best_random_warm_GBps": best, "rows": rows}), flush=True)


def copy_gbs(torch, n, stream, reps=30, rotate=1):
    """Useful GB/s of a plain device copy of n int32 (8 n bytes moved),
    back to back over `rotate` distinct source/destination pairs in turn."""
    pairs = [(torch.empty(n, dtype=torch.int32, device="cuda"),
              torch.empty(n, dtype=torch.int32, device="cuda")) for _ in range(rotate)]
    for src, dst in pairs:
        dst.copy_(src)
    a = torch.cuda.Event(enable_timing=True)
    b = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    a.record(stream)
    for r in range(reps):
        src, dst = pairs[r % rotate]
        dst.copy_(src)
    b.record(stream)
    torch.cuda.synchronize()
    return 8.0 * n * reps / (a.elapsed_time(b) * 1e-3) / 1e9


def _kernel_us(t):
    return {"tile_pass": round(t.tile_ms / t.runs * 1e3, 2),
            "level_build": round(t.levels_ms / t.runs * 1e3, 2),
            "fix_pass": round(t.fix_ms / t.runs * 1e3, 2), "launches": int(t.runs)}


def run_single(args, torch, apsep_amd):
    n = args.n
    stream = torch.cuda.current_stream()
    xs = make_inputs(torch, n)
    outs = {k: torch.empty(n, dtype=torch.int32, device="cuda") for k in CLASSES}
    torch.cuda.synchronize()
    per_reps = max(args.steps, 100)
    with apsep_amd.Scratch(args.algo, n) as s:
        t0 = time.perf_counter()
        settle(torch, s, xs, outs, stream, args.settle)
        settle_s = time.perf_counter() - t0
        per, per_warm, total_ms, rot, warm = time_solves(torch, apsep_amd, s, xs, outs,
                                                         args.steps, args.warmup, stream,
                                                         per_reps)
        cold = cold_clean(torch, s, xs, outs, n, stream)
        algo_name = {v: k for k, v in apsep_amd.ALGOS.items()}[s.algo]
    useful = 8.0 * n
    per_class = {k: round(useful / (per[k] * 1e-3) / 1e9, 1) for k in CLASSES}
    step_ms = total_ms / args.steps
    value = 3 * useful / (step_ms * 1e-3) / 1e9
    tile_ms = rot.tile_ms / rot.runs
    achieved = useful / (tile_ms * 1e-3) / 1e9
    wruns = sum(t.runs for t in warm.values())
    wtile_ms = sum(t.tile_ms for t in warm.values()) / wruns
    wachieved = useful / (wtile_ms * 1e-3) / 1e9

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
        "settle_s": round(settle_s, 2),
        "per_class_gbs": per_class,
        "per_class_us": {k: round(per[k] * 1e3, 2) for k in CLASSES},
        "per_class_regime": "3 distinct input/output pairs of the class in turn",
        "per_class_gbs_warm": {k: round(useful / (per_warm[k] * 1e-3) / 1e9, 1)
                               for k in CLASSES},
        "per_class_us_warm": {k: round(per_warm[k] * 1e3, 2) for k in CLASSES},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1),
                     "peak": PEAK_HBM_GBS, "unit": "GB/s",
                     "frac": round(achieved / PEAK_HBM_GBS, 4), "traffic": traffic,
                     "kernel": "k_tile_pass (mean over the 3 classes in turn, one "
                               "scratch/input/output per class)",
                     "algorithmic_bytes_per_launch": int(useful),
                     "warm_achieved": round(wachieved, 1),
                     "warm_frac": round(wachieved / PEAK_HBM_GBS, 4)},
        "kernel_us": _kernel_us(rot),
        "kernel_us_warm": dict(_kernel_us_from(warm), per_class={
            k: _kernel_us(t) for k, t in warm.items()}),
        "cold_clean": cold,
        "copy_gbs": round(copy_gbs(torch, n, stream, rotate=3), 1),
        "copy_gbs_warm": round(copy_gbs(torch, n, stream, rotate=1), 1),
    }
    if not args.no_cpu:
        cb, ok = cpu_baseline(torch, xs, outs, args.cpu_seconds)
        rec["cpu_baseline"] = cb
        rec["parity_checked"] = ok
        if not ok:
            print("bench.py: GPU answers differ from the oracle", file=sys.stderr)
    del xs, outs
    torch.cuda.empty_cache()
    return rec


def cold_clean(torch, s, xs, outs, n, stream, reps=10):
    """Useful GB/s per class with a clean cold cache: before each timed
    solve a 1 GiB buffer is read (summed), which evicts the 256 MB
    last-level cache without leaving dirty lines; median of `reps`.  A
    plain device copy of n int32 in the same regime beside it."""
    flush = torch.ones(1 << 28, dtype=torch.int32, device="cuda")
    acc = torch.empty((), dtype=torch.int64, device="cuda")

    def med(fn):
        evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
               for _ in range(reps)]
        torch.cuda.synchronize()
        for a, b in evs:
            torch.sum(flush, dim=0, dtype=torch.int64, out=acc)
            a.record(stream)
            fn()
            b.record(stream)
        torch.cuda.synchronize()
        return sorted(a.elapsed_time(b) for a, b in evs)[reps // 2]

    useful = 8.0 * n
    res = {}
    for k in CLASSES:
        ms = med(lambda: s.run(xs[k], outs[k], stream=stream))
        res[k] = round(useful / (ms * 1e-3) / 1e9, 1)
    src, dst = xs["random"], torch.empty(n, dtype=torch.int32, device="cuda")
    res["copy"] = round(useful / (med(lambda: dst.copy_(src)) * 1e-3) / 1e9, 1)
    s.check()
    del flush, dst
    res["frac"] = {k: round(res[k] / PEAK_HBM_GBS, 3) for k in CLASSES}
    res["regime"] = ("256 MB last-level cache evicted before each solve by reading a 1 GiB "
                     "buffer (no dirty lines left); median of %d" % reps)
    return res


def _kernel_us_from(times):
    runs = sum(t.runs for t in times.values())
    return {"tile_pass": round(sum(t.tile_ms for t in times.values()) / runs * 1e3, 2),
            "level_build": round(sum(t.levels_ms for t in times.values()) / runs * 1e3, 2),
            "fix_pass": round(sum(t.fix_ms for t in times.values()) / runs * 1e3, 2),
            "launches": int(runs)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None,
                    help="timed steps (default: 500 for the single-array headline, whose "
                         "step is ~0.2 ms; 20 otherwise)")
    ap.add_argument("--warmup", type=int, default=None,
                    help="untimed steps before the timed ones (default: 50 single, 3 otherwise)")
    ap.add_argument("--settle", type=float, default=1.0,
                    help="seconds of untimed solves before the warmup steps, so that the "
                         "GPU runs at its sustained clock")
    ap.add_argument("--n", type=int, default=1 << 25)
    ap.add_argument("--algo", default="auto")
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extras", action="store_true",
                    help="single: leave out the batch and serve sections")
    ap.add_argument("--traffic", default=os.path.join(ROOT, "profiles", "traffic.json"),
                    help="rocprofv3 FETCH/WRITE summary (bytes per launch)")
    ap.add_argument("--workload", default="single",
                    choices=["single", "batch", "serve", "sweep", "tune", "structured"],
                    help="single: the plain N=2^25 solve (headline) with the structured, "
                         "batch and serve sections; batch: the segmented CSR batch alone; "
                         "serve: the serving engine alone; sweep: sizes 2^14..2^27, "
                         "cache-warm and flushed; tune: the autotuner's choice per size "
                         "class; structured: the structure-heavy arrays at N=2^28 alone")
    ap.add_argument("--structured-n", type=int, default=1 << 28)
    ap.add_argument("--structured-reps", type=int, default=10)
    ap.add_argument("--nseg", type=int, default=8192)
    ap.add_argument("--batch-steps", type=int, default=20)
    ap.add_argument("--batch-warmup", type=int, default=3)
    ap.add_argument("--jobs", type=int, default=None,
                    help="serve: jobs per rate (default 20000 alone, 10000 in the headline)")
    ap.add_argument("--payloads", type=int, default=None,
                    help="serve: distinct jobs (default 1000)")
    ap.add_argument("--rates", default="10000,50000,150000,400000",
                    help="serve: offered jobs/s")
    args = ap.parse_args()
    if args.steps is None:
        args.steps = 500 if args.workload == "single" else 20
    if args.warmup is None:
        args.warmup = 50 if args.workload == "single" else 3
    if args.jobs is None:
        args.jobs = 20000 if args.workload == "serve" else 10000
    if args.payloads is None:
        args.payloads = 1000
    if args.gpus != 1:
        print("bench.py: this build measures one GPU only; --gpus %d is not "
              "supported (the solver does not shard across GPUs)" % args.gpus,
              file=sys.stderr)
        sys.exit(2)

    import torch
    import apsep_amd

    if args.workload == "batch":
        bench_batch(args, torch, apsep_amd)
        return
    if args.workload == "serve":
        bench_serve(args, torch, apsep_amd)
        return
    if args.workload == "sweep":
        bench_sweep(args, torch, apsep_amd)
        return
    if args.workload == "tune":
        bench_tune(args, torch, apsep_amd)
        return
    if args.workload == "structured":
        bench_structured(args, torch, apsep_amd)
        return

    rec = run_single(args, torch, apsep_amd)
    print("bench: single done (%.1f GB/s)" % rec["value"], file=sys.stderr, flush=True)
    if not args.no_extras:
        rec["structured"] = run_structured(args, torch, apsep_amd, not args.no_cpu)
        print("bench: structured done", file=sys.stderr, flush=True)
        rec["batch"] = run_batch(args, torch, apsep_amd, not args.no_cpu)
        print("bench: batch done", file=sys.stderr, flush=True)
        rec["serve"] = run_serve(args, torch, apsep_amd, not args.no_cpu)
        print("bench: serve done", file=sys.stderr, flush=True)
    print(json.dumps(rec), flush=True)


if __name__ == "__main__":
    main()
