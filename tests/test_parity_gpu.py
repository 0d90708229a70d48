"""Bit-exact parity of the gfx950 solvers against the CPU oracle.

Mirrors the reference's own tests:
  * structured known-answer cases (src/main.cu:181-222, fixture in
    tests/golden/known_answers.json);
  * seeded stress, std::mt19937 seeds 42/43/100/101 (src/main.cu:144-175,
    246-252);
  * the <, <= matrix of src/bench_spt_le.cu:93-125 (glibc rand() inputs at
    n in {2^25, 1000003, 513, 512, 511, 1}, output poisoned with 0xAB);
plus the benchmark configurations, tile/level boundary sizes, adversarial
inputs that reach the overflow and deep-search paths, and determinism.
Every call goes through the C ABI (libapsep.so via apsep_amd).
"""
import json
import os
import sys

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

ALGOS = ["wstl", "spt", "lookback"]
GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
POISON = np.int32(np.uint32(0xABABABAB).view(np.int32))


def _log(*a):
    print("[parity]", *a, file=sys.stderr, flush=True)


@pytest.fixture(scope="module")
def torch_mod():
    import torch
    assert torch.cuda.is_available()
    return torch


def solve_on(scratch, torch, host, incl=None, x_dev=None, out_dev=None):
    n = host.shape[0]
    x = torch.from_numpy(np.ascontiguousarray(host, dtype=np.int32)).cuda() \
        if x_dev is None else x_dev
    out = torch.full((max(n, 1),), int(POISON), dtype=torch.int32, device="cuda") \
        if out_dev is None else out_dev
    scratch.run(x, out, n=n, incl=incl)
    torch.cuda.synchronize()
    scratch.check()
    return out[:n].cpu().numpy()


def check(got, a, incl, what):
    exp = oracle.apsep(a, incl)
    if not np.array_equal(got, exp):
        bad = np.nonzero(got != exp)[0]
        i = int(bad[0])
        raise AssertionError("%s: %d mismatches, first i=%d val=%d got=%d exp=%d"
                             % (what, bad.size, i, a[i], got[i], exp[i]))


@pytest.mark.parametrize("algo", ALGOS)
def test_known_answers(torch_mod, algo):
    import apsep_amd as A
    with open(os.path.join(GOLDEN, "known_answers.json")) as f:
        cases = json.load(f)["cases"]
    with A.Scratch(algo, 4096) as s:
        for c in cases:
            a = np.array(c["input"], dtype=np.int32)
            for incl, key in ((False, "strict"), (True, "incl")):
                _log(algo, c["label"], key)
                got = solve_on(s, torch_mod, a, incl)
                assert got.tolist() == c[key], (algo, c["label"], key)


@pytest.mark.parametrize("algo,seed,trials,maxn", [
    ("wstl", 42, 500, 4096), ("spt", 43, 500, 4096),
    ("wstl", 100, 200, 131072), ("spt", 101, 200, 131072),
    ("lookback", 42, 500, 4096), ("lookback", 43, 500, 4096),
    ("lookback", 100, 200, 131072), ("lookback", 101, 200, 131072)])
def test_reference_stress(torch_mod, algo, seed, trials, maxn):
    import apsep_amd as A
    torch = torch_mod
    xs = torch.empty(maxn, dtype=torch.int32, device="cuda")
    outs = torch.empty(maxn, dtype=torch.int32, device="cuda")
    _log(algo, "stress seed", seed, trials, maxn)
    with A.Scratch(algo, maxn) as s:
        for t, a in enumerate(oracle.stress_inputs(seed, trials, maxn)):
            n = a.shape[0]
            xs[:n].copy_(torch.from_numpy(a))
            outs.fill_(int(POISON))
            incl = (t % 4 == 3)
            got = solve_on(s, torch, a, incl, x_dev=xs, out_dev=outs)
            check(got, a, incl, "%s stress seed=%d trial=%d n=%d" % (algo, seed, t, n))


LE_KINDS = {
    "random": lambda N: oracle.glibc_rand(1234, N),
    "dup16": lambda N: oracle.glibc_rand(99, N) % 16,
    "dup2": lambda N: oracle.glibc_rand(7, N) % 2,
    "all-equal": lambda N: np.full(N, 42, dtype=np.int32),
    "descending": lambda N: (N - np.arange(N)).astype(np.int32),
    "ascending": lambda N: np.arange(N, dtype=np.int32),
    # srand(5); h[i] = (rand() % 4 == 0) ? INT_MAX : rand()
    # (src/bench_spt_le.cu:122-125)
    "intmax": lambda N: oracle.glibc_intmax(5, N),
}


@pytest.mark.parametrize("algo", ALGOS)
@pytest.mark.parametrize("kind", list(LE_KINDS))
def test_le_matrix(torch_mod, algo, kind):
    import apsep_amd as A
    torch = torch_mod
    N = 1 << 25
    h = LE_KINDS[kind](N)
    h = np.ascontiguousarray(h, dtype=np.int32)
    x = torch.from_numpy(h).cuda()
    out = torch.empty(N, dtype=torch.int32, device="cuda")
    with A.Scratch(algo, N) as s:
        for n in (N, 1000003, 513, 512, 511, 1):
            for incl in (True, False):
                _log(algo, kind, n, "<=" if incl else "<")
                out.fill_(int(POISON))
                got = solve_on(s, torch, h[:n], incl, x_dev=x, out_dev=out)
                check(got, h[:n], incl, "%s %s n=%d incl=%d" % (algo, kind, n, incl))


@pytest.mark.parametrize("algo", ALGOS)
def test_tile_and_level_boundaries(torch_mod, algo):
    import apsep_amd as A
    torch = torch_mod
    with A.Scratch(algo, 16) as s0:
        T = s0.info.tile_elems
    # tile, 64-tile group and 4096-tile group edges
    sizes = [1, 2, 63, 64, 65, T - 1, T, T + 1, 2 * T - 1, 2 * T, 2 * T + 1,
             64 * T - 1, 64 * T, 64 * T + 1, 64 * 64 * T + 1]
    nmax = max(sizes)
    with A.Scratch(algo, nmax) as s:
        for n in sizes:
            for kind in ("random", "dup2", "ascending", "descending", "blockdesc"):
                a = oracle.fill(kind, n, seed=n)
                _log(algo, "boundary", n, kind)
                check(solve_on(s, torch, a, False), a, False, "%s n=%d %s" % (algo, n, kind))
            a = oracle.fill("dup16", n, seed=n + 1)
            check(solve_on(s, torch, a, True), a, True, "%s n=%d dup16 incl" % (algo, n))
            # equal keys across tile edges: a tile's first element answers to the
            # key before the tile under "<=" only; INT_MAX is a legal key there
            a = oracle.fill("alleq", n, seed=n)
            for incl in (False, True):
                check(solve_on(s, torch, a, incl), a, incl,
                      "%s n=%d alleq incl=%d" % (algo, n, incl))
            a = oracle.fill("intmax", n, seed=n + 2)
            check(solve_on(s, torch, a, True), a, True, "%s n=%d intmax incl" % (algo, n))


def _pairs_distinct(n, tile=8192):
    i = np.arange(n, dtype=np.int64)
    j = i % tile
    pair = i // (2 * tile)
    base = 3 * (n - pair * 2 * tile)  # later pairs lower: every pair starts a new minimum
    second = (i // tile) % 2 == 1
    return np.where(second, base + 2 * (tile - 1 - j) + 1, base + 2 * j).astype(np.int32)


@pytest.mark.parametrize("algo", ALGOS)
def test_adversarial_deep_and_overflow(torch_mod, algo):
    """Inputs whose open elements need long searches or exceed the recorded
    open list: a small first element followed by a long descent (every
    answer is 0), a global descent with periodic dips, block-descending runs,
    and many equal minima."""
    import apsep_amd as A
    torch = torch_mod
    n = 1 << 22
    i = np.arange(n, dtype=np.int64)
    cases = {
        "zero-then-descent": np.where(i == 0, 0, n - i + 5).astype(np.int32),
        "descent-with-dips": np.where(i % 9973 == 0, 7, 3 * n - i).astype(np.int32),
        "blockdesc": oracle.fill("blockdesc", n, seed=5),
        "sawtooth-desc": ((n - i) % 3000 + (i % 2)).astype(np.int32),
        "min-plateau": np.where(i % 2048 < 100, 1, 5 + (i % 7)).astype(np.int32),
        "two-level-descent": (((n - i) // 2048) * 4096 + (i % 2048)).astype(np.int32),
        # tile pairs: an ascending tile of even values, then a descending
        # tile of odd values: every open element of the second tile has its
        # own answer in the first, past its recorded spine
        "pairs-distinct-answers": _pairs_distinct(n),
        # descending runs of two equal keys after a small first key: "<"
        # leaves both keys of a pair open, "<=" answers the second in-tile
        "zero-then-paired-descent": np.where(i == 0, 0, (n - i) // 2 + 5).astype(np.int32),
    }
    with A.Scratch(algo, n) as s:
        for name, a in cases.items():
            for incl in (False, True):
                _log(algo, name, incl)
                check(solve_on(s, torch, a, incl), a, incl, "%s %s" % (algo, name))


@pytest.mark.parametrize("algo", ALGOS)
def test_largest_size(torch_mod, algo):
    """n = 2^31 - 1, the largest size the ABI accepts: the last tile is
    partial and ends at the int32 limit.  Closed forms checked on the
    device: ascending (answer i - 1), then the last tile replaced by a
    descent from 10^9 (all its elements open and beaten by the prefix min 0,
    so the fix-up rescans that tile; the answer of value v is index v - 1,
    or v under "<=")."""
    import apsep_amd as A
    torch = torch_mod
    n = (1 << 31) - 1
    x = torch.arange(n, dtype=torch.int32, device="cuda")
    out = torch.empty(n, dtype=torch.int32, device="cuda")
    t0 = (n // 8192) * 8192  # start of the last (partial) tile
    with A.Scratch(algo, n) as s:
        for incl in (False, True):
            _log(algo, "largest ascending", incl)
            out.fill_(int(POISON))
            s.run(x, out, incl=incl)
            torch.cuda.synchronize()
            s.check()
            assert out[0].item() == -1
            assert bool(torch.equal(out[1:], x[:-1])), "ascending n=2^31-1 incl=%d" % incl
        tail = 1_000_000_000 - torch.arange(n - t0, dtype=torch.int32, device="cuda")
        x[t0:] = tail
        for incl in (False, True):
            _log(algo, "largest descending last tile", incl)
            out.fill_(int(POISON))
            s.run(x, out, incl=incl)
            torch.cuda.synchronize()
            s.check()
            assert out[0].item() == -1
            head = torch.arange(1, t0, dtype=torch.int32, device="cuda")
            assert bool(torch.equal(out[1:t0], head - 1)), "head n=2^31-1 incl=%d" % incl
            assert bool(torch.equal(out[t0:], tail - (0 if incl else 1))), \
                "last tile n=2^31-1 incl=%d" % incl
            del head
    del x, out
    torch.cuda.empty_cache()


@pytest.mark.parametrize("algo", ALGOS)
def test_query_queue_overflow(torch_mod, algo):
    """More in-tile queries than a wave's queue holds (192): lanes whose 16
    values descend, so that every element but the first asks the previous
    lane.  Such waves store their outputs lane by lane; the others stage them
    through LDS and store them coalesced.  Both kinds in one tile, every
    wave of the tile, and a partial last tile."""
    import apsep_amd as A
    torch = torch_mod
    n = (1 << 21) + 777
    i = np.arange(n, dtype=np.int64)
    rng = np.random.default_rng(11)
    desc16 = (1_000_000 - (i % 16)).astype(np.int32)
    rnd = rng.integers(0, 2**31 - 1, size=n, dtype=np.int64).astype(np.int32)
    wave = (i // 1024) % 2 == 0  # 1024 elements per wave of 64 lanes
    cases = {
        "all-waves-overflow": desc16,
        "alternate-waves": np.where(wave, desc16, rnd).astype(np.int32),
        "overflow-small-values": np.where(wave, desc16 - 2_000_000, rnd).astype(np.int32),
    }
    with A.Scratch(algo, n) as s:
        for name, a in cases.items():
            for incl in (False, True):
                _log(algo, name, incl)
                check(solve_on(s, torch, a, incl), a, incl, "%s %s" % (algo, name))


@pytest.mark.parametrize("algo", ALGOS)
@pytest.mark.parametrize("n,kinds", [
    (1 << 20, ("random", "descending", "ascending")),
    (1 << 25, ("random", "descending", "ascending")),
    (1 << 28, ("dup16", "sawtooth", "blockdesc")),
])
def test_benchmark_configs(torch_mod, algo, n, kinds):
    import apsep_amd as A
    torch = torch_mod
    x = torch.empty(n, dtype=torch.int32, device="cuda")
    out = torch.empty(n, dtype=torch.int32, device="cuda")
    with A.Scratch(algo, n) as s:
        for kind in kinds:
            a = oracle.fill(kind, n, seed=2024)
            x.copy_(torch.from_numpy(a))
            out.fill_(int(POISON))
            _log(algo, "config", n, kind)
            got = solve_on(s, torch, a, False, x_dev=x, out_dev=out)
            check(got, a, False, "%s n=%d %s" % (algo, n, kind))
            # repeat on the same buffers: deterministic, nothing carried over
            got2 = solve_on(s, torch, a, False, x_dev=x, out_dev=out)
            assert np.array_equal(got, got2)


@pytest.mark.parametrize("algo", ALGOS)
def test_time_kernels_leaves_valid_output(torch_mod, algo):
    """apsep_time_kernels re-launches every kernel of the solve; the output
    afterwards is the solve's answer and every batch has a positive time."""
    import apsep_amd as A
    torch = torch_mod
    n = 3 * (1 << 20) + 17
    a = oracle.fill("random", n, seed=11)
    x = torch.from_numpy(a).cuda()
    out = torch.full((n,), int(POISON), dtype=torch.int32, device="cuda")
    with A.Scratch(algo, n) as s:
        _log(algo, "time_kernels", n)
        t = s.time_kernels(x, out, 3)
        s.check()
        assert t.runs == 3 and t.tile_ms > 0
        if algo == "wstl":
            assert t.fix_ms > 0
        check(out.cpu().numpy(), a, False, "%s time_kernels" % algo)


def test_lookback_sizes_up_and_down(torch_mod):
    """Eager look-back runs of changing sizes on one scratch: each run clears
    the group counters of the next (two sets in turn), so a large run after
    a small one must find every counter of its larger hierarchy clear.  One
    run per size, so that runs 2, 4, 6 and 8 land on a counter set that a
    run two before dirtied over more counters than the run just before
    them used (a clearing bounded by the clearing run's own size would
    leave 6, 5, 16 and 14 of their counters stale)."""
    import apsep_amd as A
    torch = torch_mod
    sizes = [1 << 22, (1 << 19) + 5, 1 << 22, (1 << 20) + 3, 1 << 23, 4099, 1 << 23,
             (1 << 19) + 5, 1 << 23]
    with A.Scratch("lookback", max(sizes)) as s:
        for k, n in enumerate(sizes):
            kind = ("random", "descending", "dup16")[k % 3]
            a = oracle.fill(kind, n, seed=70 + k)
            x = torch.from_numpy(a).cuda()
            out = torch.full((n,), int(POISON), dtype=torch.int32, device="cuda")
            _log("lookback sizes", k, n, kind)
            s.run(x, out)
            torch.cuda.synchronize()
            s.check()
            check(out.cpu().numpy(), a, False, "lookback size %d (%d) %s" % (k, n, kind))


def test_lookback_graph_replay(torch_mod):
    """The captured look-back sequence keeps no host-baked state: capture it
    once with torch.cuda.graph, then replay with new keys in the same
    buffer, eager solves on the same scratch and synchronisation between
    replays; every replay and every eager solve checked."""
    import apsep_amd as A
    torch = torch_mod
    n = (1 << 21) + 5
    x = torch.empty(n, dtype=torch.int32, device="cuda")
    out = torch.empty(n, dtype=torch.int32, device="cuda")
    with A.Scratch("lookback", n) as s:
        x.copy_(torch.from_numpy(oracle.fill("random", n, seed=1)))
        side = torch.cuda.Stream()
        side.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(side):
            s.run(x, out)  # warm-up outside the capture
        torch.cuda.current_stream().wait_stream(side)
        torch.cuda.synchronize()
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            s.run(x, out)
        for rep, kind in enumerate(["random", "descending", "dup16", "ascending", "random"]):
            a = oracle.fill(kind, n, seed=10 + rep)
            x.copy_(torch.from_numpy(a))
            out.fill_(int(POISON))
            torch.cuda.synchronize()
            _log("lookback graph replay", rep, kind)
            g.replay()
            torch.cuda.synchronize()
            s.check()
            check(out.cpu().numpy(), a, False, "lookback replay %d %s" % (rep, kind))
            # eager runs on the same scratch between replays (host-kept tags
            # and counter sets beside the replay's device-kept ones)
            for e in range(2):
                out.fill_(int(POISON))
                s.run(x, out)
                torch.cuda.synchronize()
                s.check()
                check(out.cpu().numpy(), a, False, "lookback eager %d after replay %d" % (e, rep))
