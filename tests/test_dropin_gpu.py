"""The drop-in path run for real, checked bit-exact against the oracle.

* tests/dropin/dropin_caller: a C++ program built by hipcc that includes
  only include/apsep/apsep.hpp and links libapsep.so (built twice: the
  shims as shipped, and with the SPT names forced onto the single
  persistent kernel, tests/dropin/dropin_caller_spt).  It calls
  launchWSTL / launchSPT (a device allocation per case, as the reference's
  src/main.cu:228-252 does), allocSPTScratch<int,128,4,INCL> + runSPT and
  allocWSTLScratch + runWSTL on one scratch for every case (as
  src/bench_spt_le.cu:60-91 does), and free*Scratch at the end.
* apsep_launch (the C ABI under launchWSTL / launchSPT) called directly for
  every algorithm and both predicates.
* apsep_run under graph capture: the WSTL sequence and the SPT launch are
  recorded and replayed with new keys.
"""
import os
import subprocess
import sys

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
# the shims as shipped (the SPT names run the fastest solver), and built with
# APSEP_SHIM_SPT_ALGO=APSEP_ALGO_SPT (the SPT names run the persistent kernel)
CALLERS = {"default": os.path.join(HERE, "dropin", "dropin_caller"),
           "spt": os.path.join(HERE, "dropin", "dropin_caller_spt")}
POISON = np.int32(np.uint32(0xABABABAB).view(np.int32))

# entry bits of the caller (tests/dropin/dropin_caller.cpp)
LAUNCH_WSTL, LAUNCH_SPT, LAUNCH_SPT_LE, RUN_WSTL, RUN_SPT, RUN_SPT_LE = (1 << b for b in range(6))
ENTRY_INCL = {LAUNCH_WSTL: False, LAUNCH_SPT: False, LAUNCH_SPT_LE: True,
              RUN_WSTL: False, RUN_SPT: False, RUN_SPT_LE: True}
ENTRY_NAME = {LAUNCH_WSTL: "launchWSTL", LAUNCH_SPT: "launchSPT",
              LAUNCH_SPT_LE: "launchSPT<..,true>", RUN_WSTL: "runWSTL",
              RUN_SPT: "runSPT<..,false>", RUN_SPT_LE: "runSPT<..,true>"}


def _log(*a):
    print("[dropin]", *a, file=sys.stderr, flush=True)


def run_caller(tmp_path, cases, caller, timeout=400):
    """cases: list of (label, keys, mask).  Runs the C++ caller (a key of
    CALLERS) and checks every answer it wrote against the oracle."""
    exe = CALLERS[caller]
    assert os.path.exists(exe), "%s is not built" % exe
    cin = os.path.join(str(tmp_path), "cases.bin")
    cout = os.path.join(str(tmp_path), "answers.bin")
    with open(cin, "wb") as f:
        np.array([len(cases)], dtype=np.int32).tofile(f)
        for _, keys, mask in cases:
            np.array([keys.size, mask], dtype=np.int32).tofile(f)
            np.ascontiguousarray(keys, dtype=np.int32).tofile(f)
    _log("running the C++ caller (%s) on" % caller, len(cases), "cases")
    r = subprocess.run([exe, cin, cout], capture_output=True, text=True, timeout=timeout)
    _log(r.stdout.strip().splitlines()[-1] if r.stdout.strip() else "(no output)")
    assert r.returncode == 0, "caller failed (%d):\n%s\n%s" % (r.returncode, r.stdout[-2000:],
                                                             r.stderr[-2000:])
    got_all = np.fromfile(cout, dtype=np.int32)
    pos = 0
    checked = 0
    for label, keys, mask in cases:
        n = keys.size
        exp = {False: None, True: None}
        for b in range(6):
            bit = 1 << b
            if not mask & bit:
                continue
            got = got_all[pos:pos + n]
            pos += n
            incl = ENTRY_INCL[bit]
            if exp[incl] is None:
                exp[incl] = oracle.apsep(keys, incl)
            if not np.array_equal(got, exp[incl]):
                bad = np.nonzero(got != exp[incl])[0]
                i = int(bad[0])
                raise AssertionError("%s %s n=%d: %d mismatches, first i=%d val=%d got=%d exp=%d"
                                     % (ENTRY_NAME[bit], label, n, bad.size, i, keys[i],
                                        got[i], exp[incl][i]))
            checked += 1
    assert pos == got_all.size
    return checked


def _known_cases():
    import json
    with open(os.path.join(HERE, "golden", "known_answers.json")) as f:
        return json.load(f)["cases"]


@pytest.mark.parametrize("caller", sorted(CALLERS))
def test_caller_known_answers_and_stress(tmp_path, caller):
    """src/main.cu's correctness section through the shims: the structured
    cases on every entry point, then the mt19937 stress seeds -- 42 and 100
    through launchWSTL, 43 and 101 through launchSPT, as the reference
    assigns them -- and the same trials through the run* entry points on
    one scratch."""
    cases = []
    for c in _known_cases():
        cases.append((c["label"], np.array(c["input"], dtype=np.int32), 0x3F))
    for seed, trials, maxn, mask in ((42, 500, 4096, LAUNCH_WSTL | RUN_WSTL),
                                     (43, 500, 4096, LAUNCH_SPT | LAUNCH_SPT_LE | RUN_SPT
                                      | RUN_SPT_LE),
                                     (100, 200, 131072, LAUNCH_WSTL | RUN_WSTL),
                                     (101, 200, 131072, LAUNCH_SPT | RUN_SPT | RUN_SPT_LE)):
        for t, a in enumerate(oracle.stress_inputs(seed, trials, maxn)):
            cases.append(("seed=%d trial=%d" % (seed, t), a, mask))
    n = run_caller(tmp_path, cases, caller)
    _log("checked", n, "answer arrays")


@pytest.mark.parametrize("caller", sorted(CALLERS))
def test_caller_le_matrix(tmp_path, caller):
    """src/bench_spt_le.cu's matrix through the shims: allocSPTScratch at
    N = 2^25 for each predicate, runSPT<..,INCL> on 0xAB-poisoned output at
    n in {N, 1000003, 513, 512, 511, 1} (N itself for the random and the
    INT_MAX-heavy inputs), plus runWSTL on the same cases."""
    N = 1 << 25
    kinds = {
        "random": oracle.glibc_rand(1234, N),
        "dup16": oracle.glibc_rand(99, 1000003) % 16,
        "dup2": oracle.glibc_rand(7, 1000003) % 2,
        "all-equal": np.full(1000003, 42, dtype=np.int32),
        "descending": (N - np.arange(1000003)).astype(np.int32),
        "ascending": np.arange(1000003, dtype=np.int32),
        "intmax": oracle.glibc_intmax(5, N),
    }
    mask = RUN_SPT | RUN_SPT_LE | RUN_WSTL
    cases = []
    for kind, h in kinds.items():
        h = np.ascontiguousarray(h, dtype=np.int32)
        for n in (N, 1000003, 513, 512, 511, 1):
            if n <= h.size:
                cases.append(("%s n=%d" % (kind, n), h[:n], mask))
    n = run_caller(tmp_path, cases, caller, timeout=600)
    _log("checked", n, "answer arrays")


@pytest.fixture(scope="module")
def torch_mod():
    import torch
    assert torch.cuda.is_available()
    return torch


@pytest.mark.parametrize("algo", ["wstl", "spt", "lookback", "auto"])
def test_apsep_launch_on_data(torch_mod, algo):
    """apsep_launch (allocate, solve on the default stream, synchronise,
    check the device error word, free) for each algorithm and predicate."""
    import apsep_amd as A
    torch = torch_mod
    inputs = [("random", oracle.fill("random", (1 << 20) + 3, seed=1)),
              ("descending", oracle.fill("descending", 70001)),
              ("ascending", oracle.fill("ascending", 8192 * 3)),
              ("dup16", oracle.fill("dup16", 1000003, seed=2)),
              ("one", np.array([7], dtype=np.int32))]
    for c in _known_cases():
        inputs.append((c["label"], np.array(c["input"], dtype=np.int32)))
    for label, a in inputs:
        x = torch.from_numpy(a).cuda()
        for incl in (False, True):
            out = torch.full((a.size,), int(POISON), dtype=torch.int32, device="cuda")
            _log("apsep_launch", algo, label, a.size, incl)
            A.launch(algo, x.data_ptr(), out.data_ptr(), a.size, incl)
            got = out.cpu().numpy()
            exp = oracle.apsep(a, incl)
            assert np.array_equal(got, exp), (algo, label, incl)


@pytest.mark.parametrize("algo", ["wstl", "spt"])
def test_graph_capture_and_replay(torch_mod, algo):
    """apsep_run on a WSTL or SPT scratch recorded into a graph
    (torch.cuda.graph): the run resets level 1 itself and uses the capture
    epoch (SPT also resets its barrier and discovery words with a kernel),
    so replays with new keys in the same buffer are exact, with eager runs
    on the same scratch between them (their epochs interleave with the
    captured one)."""
    import apsep_amd as A
    torch = torch_mod
    n = (1 << 22) + 12345
    x = torch.empty(n, dtype=torch.int32, device="cuda")
    out = torch.empty(n, dtype=torch.int32, device="cuda")
    x2 = torch.from_numpy(oracle.fill("random", n, seed=99)).cuda()
    out2 = torch.empty(n, dtype=torch.int32, device="cuda")
    exp2 = oracle.apsep(x2.cpu().numpy())
    with A.Scratch(algo, n) as s:
        x.copy_(torch.from_numpy(oracle.fill("random", n, seed=1)))
        side = torch.cuda.Stream()
        side.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(side):
            s.run(x, out)
        torch.cuda.current_stream().wait_stream(side)
        torch.cuda.synchronize()
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            s.run(x, out)
        kinds = ["random", "descending", "dup16", "ascending", "blockdesc", "random"]
        for rep, kind in enumerate(kinds):
            a = oracle.fill(kind, n, seed=10 + rep)
            x.copy_(torch.from_numpy(a))
            out.fill_(int(POISON))
            torch.cuda.synchronize()
            _log(algo, "graph replay", rep, kind)
            g.replay()
            torch.cuda.synchronize()
            s.check()
            got = out.cpu().numpy()
            exp = oracle.apsep(a)
            assert np.array_equal(got, exp), ("replay", rep, kind)
            # an eager run on the same scratch between replays
            out2.fill_(int(POISON))
            s.run(x2, out2)
            torch.cuda.synchronize()
            assert np.array_equal(out2.cpu().numpy(), exp2), ("eager after replay", rep)
