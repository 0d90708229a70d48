"""apsep_run_noalloc (the entry point that allocates nothing: its scratch is
the library's fixed per-GPU device array) against apsep_run_mode on a WSTL
scratch, the existing entry point with the same kernels: bit-exact on two
calls in a row, the second with a smaller input, for both predicates, and
against the oracle.  Also: sizes beyond the array are refused, and a call
recorded into a graph replays exactly (the entry keeps no state between
calls)."""
import ctypes
import sys

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

POISON = np.int32(np.uint32(0xABABABAB).view(np.int32))


def _log(*a):
    print("[no_alloc]", *a, file=sys.stderr, flush=True)


@pytest.fixture(scope="module")
def torch_mod():
    import torch
    assert torch.cuda.is_available()
    return torch


def _both(torch, A, a, incl):
    n = a.size
    x = torch.from_numpy(a).cuda()
    ref = torch.full((n,), int(POISON), dtype=torch.int32, device="cuda")
    got = torch.full((n,), int(POISON), dtype=torch.int32, device="cuda")
    with A.Scratch("wstl", n) as s:
        s.run(x, ref, incl=incl)
    A.run_noalloc(x, got, incl=incl)
    torch.cuda.synchronize()
    return got.cpu().numpy(), ref.cpu().numpy()


@pytest.mark.parametrize("incl", [False, True])
def test_noalloc_matches_run_twice(torch_mod, incl):
    import apsep_amd as A
    torch = torch_mod
    calls = [("random", A.NOALLOC_MAX_N), ("dup16", 1000003)]  # second call smaller
    for kind, n in calls:
        a = oracle.fill(kind, n, seed=n)
        _log("noalloc vs run", kind, n, "incl" if incl else "strict")
        got, ref = _both(torch, A, a, incl)
        assert np.array_equal(got, ref), (kind, n, incl)
        assert np.array_equal(got, oracle.apsep(a, incl)), (kind, n, incl)


def test_noalloc_small_after_large_and_edges(torch_mod):
    import apsep_amd as A
    torch = torch_mod
    for kind, n in [("descending", (1 << 25) + 5), ("ascending", 8193), ("random", 1),
                    ("blockdesc", 3 * 8192), ("alleq", 513)]:
        a = oracle.fill(kind, n, seed=3)
        _log("noalloc sequence", kind, n)
        got, ref = _both(torch, A, a, False)
        assert np.array_equal(got, ref), (kind, n)
        assert np.array_equal(got, oracle.apsep(a)), (kind, n)


def test_noalloc_refuses_beyond_capacity(torch_mod):
    import apsep_amd as A
    L = A.lib()
    # checked before any device access: the pointers are never dereferenced
    assert L.apsep_run_noalloc(4096, 1 << 40, A.NOALLOC_MAX_N + 1, 0, None) == A.ERR_CAPACITY
    assert L.apsep_run_noalloc(4096, 1 << 40, 10, 2, None) == A.ERR_INVALID_ARG
    assert L.apsep_run_noalloc(4100, 1 << 40, 10, 0, None) == A.ERR_ALIGN
    assert L.apsep_run_noalloc(None, None, 0, 0, None) == A.OK


def test_noalloc_graph_replay(torch_mod):
    import apsep_amd as A
    torch = torch_mod
    n = (1 << 21) + 77
    x = torch.from_numpy(oracle.fill("random", n, seed=5)).cuda()
    out = torch.empty(n, dtype=torch.int32, device="cuda")
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        A.run_noalloc(x, out)
    torch.cuda.current_stream().wait_stream(side)
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        A.run_noalloc(x, out)
    for rep, kind in enumerate(["random", "descending", "dup16", "random"]):
        a = oracle.fill(kind, n, seed=20 + rep)
        x.copy_(torch.from_numpy(a))
        out.fill_(int(POISON))
        torch.cuda.synchronize()
        _log("noalloc graph replay", rep, kind)
        g.replay()
        torch.cuda.synchronize()
        assert np.array_equal(out.cpu().numpy(), oracle.apsep(a)), (rep, kind)
        # an eager call between replays shares the array
        b = oracle.fill("ascending", 1000, seed=rep)
        xb = torch.from_numpy(b).cuda()
        ob = torch.empty(1000, dtype=torch.int32, device="cuda")
        A.run_noalloc(xb, ob)
        torch.cuda.synchronize()
        assert np.array_equal(ob.cpu().numpy(), oracle.apsep(b))
