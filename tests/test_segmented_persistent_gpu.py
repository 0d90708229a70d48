"""Segmented batches solved by one persistent launch
(apsep_batch_plan_set_persistent), bit-exact against the oracle's
per-segment monotone stack: edge lengths, every input class, both
predicates, every tile shape, thousands of tiny segments, segments whose
later tiles are open throughout, a batch deep enough for levels >= 2, graph
capture and replay with new data, switching between the persistent launch
and the kernel sequence on one plan, and two persistent batches on two
streams at once (workgroups that cannot start while the other batch holds
the GPU leave, the rest do their work).  Every test also reads the plans'
device error word (a grid barrier that timed out would raise it).
"""
import sys

import numpy as np
import pytest

import oracle

from test_segmented_gpu import EDGE_LENGTHS, POISON, _check, _keys, _offsets, _solve

pytestmark = pytest.mark.gpu


def _log(*a):
    print("[persistent]", *a, file=sys.stderr, flush=True)


@pytest.fixture(scope="module")
def torch_mod():
    import torch
    assert torch.cuda.is_available()
    return torch


@pytest.mark.parametrize("lanes", [128, 256, 512])
@pytest.mark.parametrize("incl", [False, True])
@pytest.mark.parametrize("kind", ["random", "descending", "ascending", "mixed", "intmax",
                                  "alleq"])
def test_edge_lengths(torch_mod, kind, incl, lanes):
    import apsep_amd as A
    a = _keys(kind, EDGE_LENGTHS, seed=3)
    off = _offsets(EDGE_LENGTHS)
    with A.BatchPlan(a.shape[0], len(EDGE_LENGTHS), incl) as p:
        p.set_tile_lanes(lanes)
        p.set_persistent(True)
        _log("edge", kind, incl, lanes, a.shape[0])
        _check(_solve(p, torch_mod, a, off), a, off, incl,
               "persistent edge %s incl=%d lanes=%d" % (kind, incl, lanes))
        p.check()


@pytest.mark.parametrize("lanes", [128, 256, 512])
@pytest.mark.parametrize("n", [1, 17, 2048, 2049, 8192, 70001])
def test_one_segment_small(torch_mod, n, lanes):
    """One segment of one tile or a few: the grid is a single workgroup or a
    handful, and the fix-up has zero, one or several tiles."""
    import apsep_amd as A
    a = oracle.fill("random", n, seed=n)
    off = np.array([0, n], dtype=np.int64)
    with A.BatchPlan(n, 1) as p:
        p.set_tile_lanes(lanes)
        p.set_persistent(True)
        _log("one segment", n, lanes)
        assert np.array_equal(_solve(p, torch_mod, a, off), oracle.apsep(a))
        p.check()


@pytest.mark.parametrize("lanes", [128, 256, 512])
def test_many_tiny_segments(torch_mod, lanes):
    """2^16 segments of 0..40 keys (many starts per tile and per lane): the
    setup step's segments outnumber the tiles' workgroups."""
    import apsep_amd as A
    rng = np.random.default_rng(11)
    lengths = rng.integers(0, 41, size=1 << 16)
    a = _keys("mixed", lengths, 11)
    off = _offsets(lengths)
    with A.BatchPlan(a.shape[0], len(lengths)) as p:
        p.set_tile_lanes(lanes)
        p.set_persistent(True)
        _log("tiny", lanes, len(lengths), a.shape[0])
        _check(_solve(p, torch_mod, a, off), a, off, False, "persistent tiny lanes=%d" % lanes)
        p.check()


@pytest.mark.parametrize("lanes", [128, 256, 512])
@pytest.mark.parametrize("incl", [False, True])
def test_loguniform_batch(torch_mod, incl, lanes):
    """600 segments of log-uniform lengths in [64, 2^18], mixed classes
    (more tiles than the persistent grid: every workgroup loops)."""
    import apsep_amd as A
    off = oracle.loguniform_offsets(2024, 600, 64.0, float(1 << 18))
    a = _keys("mixed", np.diff(off), 9)
    with A.BatchPlan(a.shape[0], off.shape[0] - 1, incl) as p:
        p.set_tile_lanes(lanes)
        p.set_persistent(True)
        _log("loguniform", incl, lanes, a.shape[0])
        _check(_solve(p, torch_mod, a, off), a, off, incl, "persistent loguniform")
        p.check()


@pytest.mark.parametrize("lanes", [128, 256])
def test_open_runs(torch_mod, lanes):
    """Segments whose later tiles are open throughout (a small first key,
    then a descent), block-descending segments, and descending segments
    after segments of small keys (no answer may cross a segment start)."""
    import apsep_amd as A
    rng = np.random.default_rng(17)
    parts, lengths = [], []
    for j in range(24):
        L = int(rng.integers(3 * 8192, 6 * 8192))
        i = np.arange(L, dtype=np.int64)
        kind = j % 4
        if kind == 0:
            seg = np.where(i == 0, 3, 10 * L - i)
        elif kind == 1:
            seg = oracle.fill("blockdesc", L, seed=40 + j).astype(np.int64)
        elif kind == 2:
            seg = np.full(L, 1, dtype=np.int64)
        else:
            seg = 5 * L - i
        parts.append(seg.astype(np.int32))
        lengths.append(L)
    a = np.concatenate(parts)
    off = _offsets(lengths)
    with A.BatchPlan(a.shape[0], len(lengths)) as p:
        p.set_tile_lanes(lanes)
        p.set_persistent(True)
        _log("open runs", lanes, a.shape[0])
        _check(_solve(p, torch_mod, a, off), a, off, False, "persistent open runs")
        p.check()


def test_deep_hierarchy_batch(torch_mod):
    """2^27 + 4097 keys in 64 segments on 128-lane tiles: level 1 holds more
    than 512 entries, so levels >= 2 are built between grid barriers."""
    import apsep_amd as A
    total = (1 << 27) + 4097
    rng = np.random.default_rng(6)
    cuts = np.unique(rng.integers(1, total, size=63))
    off = np.concatenate([[0], cuts, [total]]).astype(np.int64)
    a = _keys("mixed", np.diff(off), 78)
    with A.BatchPlan(total, len(off) - 1) as p:
        p.set_tile_lanes(128)
        p.set_persistent(True)
        _log("deep hierarchy", total)
        got = _solve(p, torch_mod, a, off)
        p.check()
    exp = oracle.apsep_segmented(a, off, False, threads=8)
    if not np.array_equal(got, exp):
        i = int(np.nonzero(got != exp)[0][0])
        raise AssertionError("persistent deep hierarchy: first mismatch i=%d got=%d exp=%d"
                             % (i, got[i], exp[i]))


def test_graph_replay_with_new_data(torch_mod):
    """The barrier-clearing kernel and the persistent launch captured once,
    replayed with new keys and offsets, eager runs of the same plan (kernel
    sequence and persistent) between the replays; every result checked."""
    import apsep_amd as A
    torch = torch_mod
    total, nseg = 1 << 21, 900
    x = torch.empty(total, dtype=torch.int32, device="cuda")
    o = torch.empty(nseg + 1, dtype=torch.int64, device="cuda")
    out = torch.empty(total, dtype=torch.int32, device="cuda")
    stream = torch.cuda.Stream()
    with A.BatchPlan(total, nseg) as p:
        for rep in range(5):
            rng = np.random.default_rng(200 + rep)
            cuts = np.sort(rng.integers(0, total, size=nseg - 1))
            off = np.concatenate([[0], cuts, [total]]).astype(np.int64)
            a = _keys("mixed", np.diff(off), 60 + rep)
            x.copy_(torch.from_numpy(a))
            o.copy_(torch.from_numpy(off))
            torch.cuda.synchronize()
            for graph in (True, False):
                p.set_persistent(True)
                p.use_graph(graph)
                out.fill_(int(POISON))
                torch.cuda.synchronize()
                _log("graph" if graph else "eager", rep)
                with torch.cuda.stream(stream):
                    p.run(x, o, out, total=total, stream=stream)
                stream.synchronize()
                _check(out.cpu().numpy(), a, off, False, "persistent graph=%d rep %d"
                       % (graph, rep))
            # the kernel sequence on the same plan between replays
            p.set_persistent(False)
            out.fill_(int(POISON))
            torch.cuda.synchronize()
            with torch.cuda.stream(stream):
                p.run(x, o, out, total=total, stream=stream)
            stream.synchronize()
            _check(out.cpu().numpy(), a, off, False, "sequence rep %d" % rep)
        p.check()


def test_two_batches_on_two_streams(torch_mod):
    """Two plans' persistent launches queued on two streams, eight rounds:
    each launch's grid is sized to fill the GPU alone, so workgroups of one
    may find no room while the other runs (they leave; the participants
    share the work).  Every answer checked, no device error."""
    import apsep_amd as A
    torch = torch_mod
    batches = []
    for seed in (31, 32):
        off = oracle.loguniform_offsets(seed, 400, 64.0, float(1 << 17))
        a = _keys("mixed", np.diff(off), seed)
        batches.append((a, off))
    streams = [torch.cuda.Stream(), torch.cuda.Stream()]
    plans = [A.BatchPlan(a.shape[0], off.shape[0] - 1) for a, off in batches]
    try:
        xs, os_, outs = [], [], []
        for (a, off), p in zip(batches, plans):
            p.set_tile_lanes(128)
            p.set_persistent(True)
            xs.append(torch.from_numpy(a).cuda())
            os_.append(torch.from_numpy(off).cuda())
            outs.append(torch.full((a.shape[0],), int(POISON), dtype=torch.int32,
                                   device="cuda"))
        torch.cuda.synchronize()
        for rnd in range(8):
            _log("two streams round", rnd)
            for k in range(2):
                with torch.cuda.stream(streams[k]):
                    plans[k].run(xs[k], os_[k], outs[k], stream=streams[k])
            torch.cuda.synchronize()
            for k in range(2):
                a, off = batches[k]
                _check(outs[k].cpu().numpy(), a, off, False, "two streams round %d plan %d"
                       % (rnd, k))
                outs[k].fill_(int(POISON))
        for p in plans:
            p.check()
    finally:
        for p in plans:
            p.close()
