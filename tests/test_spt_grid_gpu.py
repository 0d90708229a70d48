"""The persistent solve (SPT) on grids of any size, bit-exact against the
oracle.

The persistent kernel joins its workgroups through eight slot counters and
gives slot c the tiles of residue c; a residue whose slot has no
participant is solved by all participants after their own, and a grid with
more workgroups than fit at once leaves the late ones out.  Here the grid
is set explicitly (apsep_scratch_set_persistent_blocks): a single
workgroup, fewer workgroups than slots, uneven slots, and twice the
co-resident grid on an array with more tiles than that, each on every
input class with both predicates.  Every call goes through the C ABI.
"""
import sys

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

TILE = 8192


def _log(*a):
    print("[spt-grid]", *a, file=sys.stderr, flush=True)


@pytest.fixture(scope="module")
def torch_mod():
    import torch
    assert torch.cuda.is_available()
    return torch


def _solve_check(torch, s, a, incl, what):
    n = a.size
    x = torch.from_numpy(a).cuda()
    out = torch.full((n,), -7, dtype=torch.int32, device="cuda")
    s.run(x, out, incl=incl)
    torch.cuda.synchronize()
    s.check()
    got = out.cpu().numpy()
    exp = oracle.apsep(a, incl)
    if not np.array_equal(got, exp):
        bad = np.nonzero(got != exp)[0]
        i = int(bad[0])
        raise AssertionError("%s: %d mismatches, first i=%d got=%d exp=%d"
                             % (what, bad.size, i, got[i], exp[i]))


@pytest.mark.parametrize("blocks", [1, 3, 7, 9, 100])
def test_small_grids(torch_mod, blocks):
    import apsep_amd as A
    sizes = [TILE * 5 + 17, TILE * 40, (1 << 21) + 3]
    with A.Scratch("spt", max(sizes)) as s:
        s.set_persistent_blocks(blocks)
        for n in sizes:
            for kind in ("random", "descending", "ascending", "dup16"):
                a = oracle.fill(kind, n, seed=n % 1000 + blocks)
                for incl in (False, True):
                    _log("blocks", blocks, "n", n, kind, "incl" if incl else "strict")
                    _solve_check(torch_mod, s, a, incl, "blocks=%d n=%d %s incl=%s"
                                 % (blocks, n, kind, incl))


def test_grid_beyond_coresident(torch_mod):
    """Twice the default grid on 2^24 keys (2048 tiles): the workgroups that
    cannot start while the first ones run arrive after the count is closed
    and leave; the participants solve every tile."""
    import apsep_amd as A
    n = 1 << 24
    with A.Scratch("spt", n) as s:
        blocks = 2 * s.info.persistent_blocks
        s.set_persistent_blocks(blocks)
        for kind in ("random", "descending"):
            a = oracle.fill(kind, n, seed=5)
            for incl in (False, True):
                _log("blocks", blocks, "n", n, kind, "incl" if incl else "strict")
                _solve_check(torch_mod, s, a, incl, "blocks=%d %s incl=%s" % (blocks, kind, incl))
        s.set_persistent_blocks(0)  # back to the default grid
        a = oracle.fill("random", n, seed=6)
        _solve_check(torch_mod, s, a, False, "default grid after")


def test_set_blocks_rejects(torch_mod):
    import apsep_amd as A
    with A.Scratch("spt", TILE) as s:
        with pytest.raises(A.ApsepError):
            s.set_persistent_blocks(-1)
    with A.Scratch("wstl", TILE) as s:
        with pytest.raises(A.ApsepError):
            s.set_persistent_blocks(4)


def test_cooperative_launch(torch_mod):
    """apsep_scratch_set_cooperative: the persistent kernel launched with
    hipLaunchCooperativeKernel (as the reference's runSPT launches with
    cudaLaunchCooperativeKernel), on the default grid and on grids of 7 and
    twice the co-resident size (clamped to what fits at once), alternating
    with normal launches on the same scratch; every class, both predicates.
    A launch the runtime refuses runs normally and is counted."""
    import apsep_amd as A
    n = (1 << 24) + 5
    with A.Scratch("spt", n) as s:
        for blocks in (0, 7, 2 * s.info.persistent_blocks):
            s.set_persistent_blocks(blocks)
            for kind in ("random", "descending", "ascending"):
                a = oracle.fill(kind, n, seed=40 + blocks % 97)
                for incl in (False, True):
                    for coop in (True, False):
                        s.set_cooperative(coop)
                        _log("cooperative" if coop else "normal", "blocks", blocks, kind,
                             "incl" if incl else "strict")
                        _solve_check(torch_mod, s, a, incl, "coop=%d blocks=%d %s incl=%s"
                                     % (coop, blocks, kind, incl))
        _log("cooperative refusals", s.cooperative_refusals)
        assert s.cooperative_refusals >= 0
    with A.Scratch("wstl", TILE) as s:
        with pytest.raises(A.ApsepError):
            s.set_cooperative(True)
