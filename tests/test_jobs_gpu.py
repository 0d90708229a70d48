"""In-place job batches (apsep_run_jobs): many independent arrays solved in
one launch sequence, each read where it lies and answered into its own
buffer, bit-exact against the oracle's per-array stack.  Covers empty and
one-key jobs, jobs at lane / chunk / tile edges and spanning many tiles,
every input class, both predicates, both tile shapes, tiles that no job
covers, graph replay with a changed job table, and the 8192-array batched
configuration checked against the segmented solve of the same arrays."""
import sys

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

POISON = np.int32(np.uint32(0xABABABAB).view(np.int32))
KINDS = ("random", "descending", "ascending", "dup16", "dup2", "alleq", "intmax", "blockdesc")


def _log(*a):
    print("[jobs]", *a, file=sys.stderr, flush=True)


@pytest.fixture(scope="module")
def torch_mod():
    import torch
    assert torch.cuda.is_available()
    return torch


def _make(torch, lengths, seed):
    rng = np.random.default_rng(seed)
    host, xs, outs = [], [], []
    for j, n in enumerate(lengths):
        a = oracle.fill(KINDS[rng.integers(len(KINDS))], int(n), seed=seed * 7919 + j) \
            if n else np.zeros(0, np.int32)
        host.append(a)
        # separate allocations, as independent requests would have them
        xs.append(torch.from_numpy(a).cuda() if n else torch.zeros(4, dtype=torch.int32,
                                                                    device="cuda")[:0])
        outs.append(torch.full((max(int(n), 1),), int(POISON), dtype=torch.int32, device="cuda"))
    return host, xs, outs


def _check(host, outs, incl, what):
    import torch
    torch.cuda.synchronize()
    for k, a in enumerate(host):
        got = outs[k][:a.size].cpu().numpy()
        exp = oracle.apsep(a, incl)
        if not np.array_equal(got, exp):
            i = int(np.nonzero(got != exp)[0][0])
            raise AssertionError("%s: job %d (n=%d) i=%d got=%d exp=%d"
                                 % (what, k, a.size, i, got[i], exp[i]))


def _edge_lengths(T):
    return [0, 1, 2, 3, 15, 16, 17, 63, 64, 65, T - 1, T, T + 1, 0, 2 * T - 1, 2 * T,
            2 * T + 1, 64 * T + 5, 1, 0, 3 * T + 64, 100000, 7]


@pytest.mark.parametrize("lanes", [128, 256, 512])
@pytest.mark.parametrize("incl", [False, True])
def test_jobs_edges_and_classes(torch_mod, lanes, incl):
    import apsep_amd as A
    torch = torch_mod
    T = 16 * lanes
    lengths = _edge_lengths(T)
    host, xs, outs = _make(torch, lengths, seed=3 + lanes)
    total = sum(lengths) + len(lengths) * T
    with A.BatchPlan(total, len(lengths), incl=incl) as p:
        p.set_tile_lanes(lanes)
        assert p.tile_elems == T
        _log("edges", lanes, incl, len(lengths))
        p.run_jobs(xs, outs)
        _check(host, outs, incl, "edges lanes=%d incl=%d" % (lanes, incl))


@pytest.mark.parametrize("lanes", [128, 256, 512])
def test_jobs_many_small_and_log_uniform(torch_mod, lanes):
    import apsep_amd as A
    torch = torch_mod
    rng = np.random.default_rng(17)
    small = list(rng.integers(0, 40, size=3000))
    big = list(np.floor(np.exp(rng.uniform(np.log(64), np.log(1 << 18), size=300))).astype(int))
    lengths = small + big
    rng.shuffle(lengths)
    host, xs, outs = _make(torch, lengths, seed=5)
    T = 16 * lanes
    with A.BatchPlan(int(sum(lengths)) + len(lengths) * T, len(lengths)) as p:
        p.set_tile_lanes(lanes)
        _log("mixed", lanes, len(lengths), int(sum(lengths)))
        p.run_jobs(xs, outs)
        _check(host, outs, False, "mixed lanes=%d" % lanes)


def test_jobs_uncovered_tiles_and_capacity(torch_mod):
    """Tiles beyond the last job (a larger tile total than the jobs need)
    are solved as empty; a total above the plan's tiles is refused."""
    import apsep_amd as A
    torch = torch_mod
    lengths = [5000, 1, 70000]
    host, xs, outs = _make(torch, lengths, seed=9)
    with A.BatchPlan(1 << 20, 8) as p:
        table, tiles = p.job_table(xs, outs)
        p.run_jobs(xs, outs, table=table, tiles=tiles + 7)
        _check(host, outs, False, "uncovered tiles")
        with pytest.raises(A.ApsepError) as e:
            p.run_jobs(xs, outs, table=table, tiles=p.job_tiles + 1)
        assert e.value.status == A.ERR_CAPACITY


def test_jobs_graph_replay_with_new_table(torch_mod):
    """With graphs on, the sequence is recorded once; replays with new keys
    and a rewritten job table (same size) are exact, with eager runs of the
    same plan between them."""
    import apsep_amd as A
    torch = torch_mod
    rng = np.random.default_rng(23)
    stream = torch.cuda.Stream()
    with A.BatchPlan(1 << 22, 64) as p, torch.cuda.stream(stream):
        p.use_graph(True)
        table = None
        for rep in range(4):
            lengths = [int(x) for x in rng.integers(0, 60000, size=40)]
            host, xs, outs = _make(torch, lengths, seed=100 + rep)
            t, tiles = p.job_table(xs, outs)
            if table is None:
                table = t
                tiles0 = p.job_tiles  # one recorded size for every replay
            else:
                table.copy_(t)
            torch.cuda.synchronize()
            _log("graph replay", rep)
            p.run_jobs(xs, outs, stream=stream, table=table, tiles=tiles0)
            _check(host, outs, False, "graph replay %d" % rep)
            # the null stream takes the eager path of the same plan
            h2, x2, o2 = _make(torch, [1000, 0, 9000], seed=200 + rep)
            torch.cuda.synchronize()
            p.run_jobs(x2, o2, stream=torch.cuda.default_stream())
            _check(h2, o2, False, "eager between replays %d" % rep)


def test_jobs_match_segmented_on_batched_configuration(torch_mod):
    """The 8192-array batched configuration (log-uniform lengths 64..2^20,
    class per array uniform over random / descending / ascending) solved as
    in-place jobs, each array in its own allocation: identical to the
    segmented solve of the concatenation and to the oracle."""
    import os
    import apsep_amd as A
    torch = torch_mod
    rng = np.random.default_rng(2024)
    nseg = 8192
    lengths = np.floor(np.exp(rng.uniform(np.log(64), np.log(1 << 20), size=nseg))).astype(np.int64)
    cls = rng.integers(0, 3, size=nseg)
    off = np.zeros(nseg + 1, dtype=np.int64)
    np.cumsum(lengths, out=off[1:])
    a = np.empty(int(off[-1]), dtype=np.int32)
    kinds = ("random", "descending", "ascending")
    for k in range(nseg):
        oracle.fill(kinds[cls[k]], int(lengths[k]), seed=k, out=a[off[k]:off[k + 1]])
    threads = max(1, min(16, len(os.sched_getaffinity(0))))
    exp = oracle.apsep_segmented(a, off, False, threads=threads)
    # each array in its own 16-byte aligned slot of one buffer
    slot = np.zeros(nseg + 1, dtype=np.int64)
    np.cumsum((lengths + 3) // 4 * 4, out=slot[1:])
    x = torch.zeros(int(slot[-1]), dtype=torch.int32, device="cuda")
    for k in range(nseg):
        x[slot[k]:slot[k] + lengths[k]] = torch.from_numpy(a[off[k]:off[k + 1]])
    xs = [x[slot[k]:slot[k] + lengths[k]] for k in range(nseg)]
    outs = [torch.empty(int(lengths[k]), dtype=torch.int32, device="cuda") for k in range(nseg)]
    with A.BatchPlan(int(off[-1]) + nseg * 8192, nseg) as p:
        for lanes in (256, 512):
            p.set_tile_lanes(lanes)
            for o in outs:
                o.fill_(int(POISON))
            _log("batched configuration as jobs", lanes)
            p.run_jobs(xs, outs)
            torch.cuda.synchronize()
            got = torch.cat(outs).cpu().numpy()
            assert np.array_equal(got, exp), "jobs lanes=%d differ from the oracle" % lanes
