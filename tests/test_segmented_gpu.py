"""Segmented (CSR batch) solves through the C ABI, bit-exact against the
oracle's per-segment monotone stack (answers relative to each segment
start).  Covers empty and one-element segments, segments that start inside
a lane, at lane/chunk/tile edges and span many tiles, every input class,
both predicates, every tile shape (128, 256 and 512 lanes), the autotuner's
choice, and hipGraph capture/replay with new data between replays.
"""
import sys

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

POISON = np.int32(np.uint32(0xABABABAB).view(np.int32))


def _log(*a):
    print("[segmented]", *a, file=sys.stderr, flush=True)


@pytest.fixture(scope="module")
def torch_mod():
    import torch
    assert torch.cuda.is_available()
    return torch


def _offsets(lengths):
    off = np.zeros(len(lengths) + 1, dtype=np.int64)
    np.cumsum(np.asarray(lengths, dtype=np.int64), out=off[1:])
    return off


def _keys(kind, lengths, seed):
    """Concatenated keys, one class per segment (or `kind` for all)."""
    parts = []
    rng = np.random.default_rng(seed)
    kinds = ("random", "descending", "ascending", "dup16", "dup2")
    for j, L in enumerate(lengths):
        k = kind if kind != "mixed" else kinds[rng.integers(len(kinds))]
        parts.append(oracle.fill(k, int(L), seed=seed + j) if L else np.zeros(0, np.int32))
    return np.concatenate(parts).astype(np.int32) if parts else np.zeros(0, np.int32)


def _solve(plan, torch, a, off, stream=None):
    x = torch.from_numpy(a).cuda()
    o = torch.from_numpy(off).cuda()
    out = torch.full((max(a.shape[0], 1),), int(POISON), dtype=torch.int32, device="cuda")
    plan.run(x, o, out, total=a.shape[0], stream=stream)
    torch.cuda.synchronize()
    return out[:a.shape[0]].cpu().numpy()


def _check(got, a, off, incl, what):
    exp = oracle.apsep_segmented(a, off, incl)
    if not np.array_equal(got, exp):
        bad = np.nonzero(got != exp)[0]
        i = int(bad[0])
        k = int(np.searchsorted(off, i, side="right") - 1)
        raise AssertionError("%s: %d mismatches, first i=%d (segment %d at %d) got=%d exp=%d"
                             % (what, bad.size, i, k, off[k], got[i], exp[i]))


EDGE_LENGTHS = [0, 1, 2, 15, 16, 17, 0, 63, 64, 65, 4095, 4096, 4097, 1, 0, 8191, 12289,
                3, 70000, 5, 0, 4096 * 70 + 3, 7, 1]


@pytest.mark.parametrize("lanes", [128, 256, 512])
@pytest.mark.parametrize("incl", [False, True])
@pytest.mark.parametrize("kind", ["random", "descending", "ascending", "dup2", "mixed",
                                  "intmax", "alleq"])
def test_edge_lengths(torch_mod, kind, incl, lanes):
    import apsep_amd as A
    lengths = EDGE_LENGTHS
    a = _keys(kind, lengths, seed=3)
    off = _offsets(lengths)
    with A.BatchPlan(a.shape[0], len(lengths), incl) as p:
        p.set_tile_lanes(lanes)
        assert p.tile_lanes == lanes
        _log(kind, incl, lanes, a.shape[0])
        _check(_solve(p, torch_mod, a, off), a, off, incl,
               "edge %s incl=%d lanes=%d" % (kind, incl, lanes))


@pytest.mark.parametrize("lanes", [128, 256, 512])
@pytest.mark.parametrize("incl", [False, True])
def test_segments_with_open_runs(torch_mod, incl, lanes):
    """Segments spanning several tiles whose later tiles are open throughout
    (more open elements than a tile records): a small first key then a
    descent (every answer is the segment's first key), block-descending
    runs, and descending segments after a segment of small keys (no answer
    may cross into the earlier segment)."""
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
            seg = np.full(L, 1, dtype=np.int64)  # small keys ...
        else:
            seg = 5 * L - i  # ... then a descent: every answer is -1
        parts.append(seg.astype(np.int32))
        lengths.append(L)
    a = np.concatenate(parts)
    off = _offsets(lengths)
    with A.BatchPlan(a.shape[0], len(lengths), incl) as p:
        p.set_tile_lanes(lanes)
        _log("open runs", incl, lanes, a.shape[0])
        _check(_solve(p, torch_mod, a, off), a, off, incl,
               "open runs incl=%d lanes=%d" % (incl, lanes))


@pytest.mark.parametrize("lanes", [128, 256, 512])
@pytest.mark.parametrize("seed", [1, 2, 3])
def test_many_small_segments(torch_mod, seed, lanes):
    """Thousands of segments shorter than a lane / a tile (many starts per
    tile, several per lane)."""
    import apsep_amd as A
    rng = np.random.default_rng(seed)
    lengths = rng.integers(0, 40, size=5000)
    a = _keys("mixed", lengths, seed)
    off = _offsets(lengths)
    with A.BatchPlan(a.shape[0], len(lengths)) as p:
        p.set_tile_lanes(lanes)
        _log("small", seed, lanes, len(lengths), a.shape[0])
        _check(_solve(p, torch_mod, a, off), a, off, False,
               "small seed=%d lanes=%d" % (seed, lanes))


@pytest.mark.parametrize("lanes", [128, 256, 512])
def test_loguniform_batch(torch_mod, lanes):
    """A scaled-down batched configuration: log-uniform lengths in
    [64, 2^18], class uniform over random/descending/ascending."""
    import apsep_amd as A
    off = oracle.loguniform_offsets(2024, 600, 64.0, float(1 << 18))
    lengths = np.diff(off)
    a = _keys("mixed", lengths, 9)
    with A.BatchPlan(a.shape[0], len(lengths)) as p:
        p.set_tile_lanes(lanes)
        _log("loguniform", lanes, len(lengths), a.shape[0])
        _check(_solve(p, torch_mod, a, off), a, off, False, "loguniform")
        # the same plan again, smaller batch (capacity is an upper bound)
        off2 = off[:301]
        a2 = a[:int(off2[-1])]
        _check(_solve(p, torch_mod, a2, off2), a2, off2, False, "loguniform prefix")


def test_tune_picks_a_shape_and_switching_shapes_keeps_answers(torch_mod):
    """The autotuner times both tile shapes (direct and graph) on the batch
    and keeps the fastest; its output is the answer.  Switching the shape of
    a graph plan afterwards drops the recorded graph, and both shapes give
    the same bits."""
    import apsep_amd as A
    torch = torch_mod
    off = oracle.loguniform_offsets(77, 400, 64.0, float(1 << 17))
    lengths = np.diff(off)
    a = _keys("mixed", lengths, 21)
    total = a.shape[0]
    x = torch.from_numpy(a).cuda()
    o = torch.from_numpy(off).cuda()
    out = torch.full((total,), int(POISON), dtype=torch.int32, device="cuda")
    stream = torch.cuda.Stream()
    with A.BatchPlan(total, len(lengths)) as p:
        _log("tune", len(lengths), total)
        with torch.cuda.stream(stream):
            e, g = p.tune(x, o, out, total=total, stream=stream, reps=2)
        stream.synchronize()
        assert e > 0 and g > 0
        chosen = p.tile_lanes
        assert chosen in (128, 256, 512)
        _check(out.cpu().numpy(), a, off, False, "after tune (lanes=%d)" % chosen)
        p.use_graph(True)
        for lanes in (512, 128, 256, 512):
            p.set_tile_lanes(lanes)
            out.fill_(int(POISON))
            torch.cuda.synchronize()
            with torch.cuda.stream(stream):
                p.run(x, o, out, total=total, stream=stream)
            stream.synchronize()
            _check(out.cpu().numpy(), a, off, False, "graph plan lanes=%d" % lanes)


@pytest.mark.parametrize("lanes", [128, 256, 512])
@pytest.mark.parametrize("cap", [3 * 8192 + 1, 5 * 8192 + 4095, 7 * 4096 + 17])
def test_capacity_off_tile_multiples(torch_mod, cap, lanes):
    """Plans whose capacity is not a multiple of either tile, filled to
    capacity: the last partial tile of the larger shape writes all of its
    chunk mins, which the plan's scratch (sized for the smaller tile) must
    hold.  An overrun would corrupt the tile counts stored after them."""
    import apsep_amd as A
    lengths = [cap // 3, cap // 3, cap - 2 * (cap // 3)]
    for kind in ("random", "ascending", "mixed"):
        a = _keys(kind, lengths, seed=cap % 97)
        off = _offsets(lengths)
        with A.BatchPlan(cap, len(lengths)) as p:
            p.set_tile_lanes(lanes)
            _log("capacity", cap, lanes, kind)
            _check(_solve(p, torch_mod, a, off), a, off, False,
                   "cap=%d lanes=%d %s" % (cap, lanes, kind))


@pytest.mark.parametrize("lanes", [128, 256, 512])
def test_deep_hierarchy_batch(torch_mod, lanes):
    """2^28 + 12345 keys in 64 segments: level 1 holds more than 512 entries
    with either tile shape, so levels >= 2 are built (one launch per level)
    and the fix-up's prefix-mins and searches walk them."""
    import apsep_amd as A
    total = (1 << 28) + 12345
    rng = np.random.default_rng(5)
    cuts = np.unique(rng.integers(1, total, size=63))
    off = np.concatenate([[0], cuts, [total]]).astype(np.int64)
    a = _keys("mixed", np.diff(off), 77)
    with A.BatchPlan(total, len(off) - 1) as p:
        p.set_tile_lanes(lanes)
        _log("deep hierarchy", lanes, total)
        got = _solve(p, torch_mod, a, off)
    exp = oracle.apsep_segmented(a, off, False, threads=8)
    if not np.array_equal(got, exp):
        i = int(np.nonzero(got != exp)[0][0])
        raise AssertionError("deep hierarchy lanes=%d: first mismatch i=%d got=%d exp=%d"
                             % (lanes, i, got[i], exp[i]))


def test_one_segment_equals_plain_solve(torch_mod):
    import apsep_amd as A
    n = (1 << 22) + 77
    a = oracle.fill("random", n, seed=5)
    off = np.array([0, n], dtype=np.int64)
    with A.BatchPlan(n, 1) as p:
        got = _solve(p, torch_mod, a, off)
    assert np.array_equal(got, oracle.apsep(a))


def test_graph_replay_with_new_data(torch_mod):
    """Capture once, replay with new keys and new offsets in the same
    buffers, eager work and synchronisation between replays; every replay
    checked."""
    import apsep_amd as A
    torch = torch_mod
    total, nseg = 1 << 21, 900
    x = torch.empty(total, dtype=torch.int32, device="cuda")
    o = torch.empty(nseg + 1, dtype=torch.int64, device="cuda")
    out = torch.empty(total, dtype=torch.int32, device="cuda")
    stream = torch.cuda.Stream()
    with A.BatchPlan(total, nseg, graph=True) as p:
        for rep in range(5):
            rng = np.random.default_rng(100 + rep)
            cuts = np.sort(rng.integers(0, total, size=nseg - 1))
            off = np.concatenate([[0], cuts, [total]]).astype(np.int64)
            a = _keys("mixed", np.diff(off), 50 + rep)
            x.copy_(torch.from_numpy(a))
            o.copy_(torch.from_numpy(off))
            out.fill_(int(POISON))
            torch.cuda.synchronize()
            _log("graph replay", rep)
            with torch.cuda.stream(stream):
                p.run(x, o, out, total=total, stream=stream)
            stream.synchronize()
            _check(out.cpu().numpy(), a, off, False, "graph replay %d" % rep)


def _c4_batch(seed=2024):
    """The batched configuration at its full size: 8192 segments, lengths
    floor(exp(U(ln 64, ln 2^20))) (numpy seed), class per segment uniform
    over random / strictly descending / strictly ascending, int64 CSR
    offsets; about 8.7e8 keys."""
    rng = np.random.default_rng(seed)
    nseg = 8192
    lengths = np.floor(np.exp(rng.uniform(np.log(64), np.log(1 << 20), size=nseg))).astype(np.int64)
    cls = rng.integers(0, 3, size=nseg)
    off = _offsets(lengths)
    a = np.empty(int(off[-1]), dtype=np.int32)
    kinds = ("random", "descending", "ascending")
    for k in range(nseg):
        oracle.fill(kinds[cls[k]], int(lengths[k]), seed=seed + k, out=a[off[k]:off[k + 1]])
    return a, off


def test_c4_full_batch_eager_and_graph(torch_mod):
    """The full batched configuration in one plan, checked bit-exact over
    the whole batch against the oracle's per-segment stack: direct launches
    in both tile shapes, then the autotuner's choice replayed as a hipGraph
    (twice, with the answers poisoned between replays)."""
    import os
    import apsep_amd as A
    torch = torch_mod
    a, off = _c4_batch()
    total, nseg = a.shape[0], off.shape[0] - 1
    _log("C4 batch", nseg, "segments,", total, "keys")
    threads = max(1, min(16, len(os.sched_getaffinity(0))))
    exp = oracle.apsep_segmented(a, off, False, threads=threads)
    x = torch.from_numpy(a).cuda()
    o = torch.from_numpy(off).cuda()
    out = torch.empty(total, dtype=torch.int32, device="cuda")

    def check(what):
        torch.cuda.synchronize()
        got = out.cpu().numpy()
        if not np.array_equal(got, exp):
            bad = np.nonzero(got != exp)[0]
            i = int(bad[0])
            k = int(np.searchsorted(off, i, side="right") - 1)
            raise AssertionError("C4 %s: %d mismatches, first i=%d (segment %d) got=%d exp=%d"
                                 % (what, bad.size, i, k, got[i], exp[i]))

    stream = torch.cuda.Stream()
    with A.BatchPlan(total, nseg) as p, torch.cuda.stream(stream):
        for lanes in (256, 512):
            p.set_tile_lanes(lanes)
            out.fill_(int(POISON))
            _log("C4 eager", lanes)
            p.run(x, o, out, stream=stream)
            check("eager lanes=%d" % lanes)
        # one persistent launch (apsep_batch_plan_set_persistent)
        p.set_persistent(True)
        out.fill_(int(POISON))
        _log("C4 persistent", p.tile_lanes)
        p.run(x, o, out, stream=stream)
        check("persistent lanes=%d" % p.tile_lanes)
        p.check()
        p.set_persistent(False)
        p.tune(x, o, out, stream=stream)
        check("autotuned")
        p.use_graph(True)
        for rep in range(2):
            out.fill_(int(POISON))
            _log("C4 graph replay", rep, "lanes", p.tile_lanes)
            p.run(x, o, out, stream=stream)
            check("graph replay %d lanes=%d" % (rep, p.tile_lanes))
    del x, out
    torch.cuda.empty_cache()


def test_tuned_size_classes(torch_mod):
    """The autotuner records its choice per size class; later batches of a
    tuned class run with that class's shape and mode, batches of an untuned
    class with the plan's current settings, and every answer is exact."""
    import apsep_amd as A
    torch = torch_mod
    stream = torch.cuda.Stream()
    batches = []
    for seed, nseg, hi in ((1, 40, 1 << 12), (2, 300, 1 << 17), (3, 200, 1 << 20)):
        off = oracle.loguniform_offsets(seed, nseg, 64.0, float(hi))
        a = _keys("mixed", np.diff(off), seed)
        batches.append((a, off))
    cap = max(a.shape[0] for a, _ in batches)
    with A.BatchPlan(cap, 300) as p, torch.cuda.stream(stream):
        for a, off in batches[:2]:  # tune two classes
            x = torch.from_numpy(a).cuda()
            o = torch.from_numpy(off).cuda()
            out = torch.empty_like(x)
            p.tune(x, o, out, stream=stream)
            torch.cuda.synchronize()
            _check(out.cpu().numpy(), a, off, False, "tune")
        tuned = [p.tuned(c) for c in range(32)]
        classes = [c for c in range(32) if tuned[c][0]]
        _log("tuned classes", [(c, tuned[c]) for c in classes])
        assert len(classes) == 2 and all(tuned[c][0] in (128, 256, 512) for c in classes)
        _log("tuned sequences", [(c, p.tuned_persistent(c)) for c in classes])
        for rep in range(2):
            for i, (a, off) in enumerate(batches):  # the third class is untuned
                _log("class run", rep, i, a.shape[0])
                _check(_solve(p, torch, a, off, stream=stream), a, off, False,
                       "class run %d batch %d" % (rep, i))
