"""Pins the CPU oracle (oracle/apsep_oracle.c) before anything is checked
against it.  CPU only.

Each check is chosen so a plausible mistake in the oracle (a dropped pop,
<= for <, an off-by-one, a wrong generator constant) fails at least one of:
the reference's own known-answer cases (src/main.cu:181-222), closed forms,
the O(n^2) definition scan, the clause-wise verifier, the C++ standard's
required mt19937 value and libc's own rand().
"""
import ctypes
import json
import os

import numpy as np
import pytest

import oracle

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def _known():
    with open(os.path.join(GOLDEN, "known_answers.json")) as f:
        return json.load(f)["cases"]


# ---------------------------------------------------------------- known answers

def test_known_answers_fixture_matches_oracle():
    for c in _known():
        a = np.array(c["input"], dtype=np.int32)
        assert oracle.apsep(a).tolist() == c["strict"], c["label"]
        assert oracle.apsep(a, incl=True).tolist() == c["incl"], c["label"]


def test_known_answers_closed_forms():
    """The fixture's expected outputs, re-derived without any solver."""
    cases = {c["label"]: c for c in _known()}
    assert cases["single element"]["strict"] == [-1]
    assert cases["two ascending"]["strict"] == [-1, 0]
    assert cases["two descending"]["strict"] == [-1, -1]
    assert cases["all equal"]["strict"] == [-1] * 5
    assert cases["all equal"]["incl"] == [-1, 0, 1, 2, 3]
    assert cases["ascending"]["strict"] == [-1] + list(range(7))
    assert cases["descending"]["strict"] == [-1] * 8
    # hand-derived (src/main.cu:191)
    assert cases["alternating"]["strict"] == [-1, -1, 1, -1, 3, 4, 3, 6]
    assert cases["exact block (B=512)"]["strict"] == [-1] + list(range(511))
    # (i%7)*13 and i%17: -1 at the period start, else the previous index
    assert cases["two full blocks"]["strict"] == [-1 if i % 7 == 0 else i - 1
                                                  for i in range(1024)]
    assert cases["non-power-of-two"]["strict"] == [-1 if i % 17 == 0 else i - 1
                                                   for i in range(525)]
    assert cases["ties"]["strict"] == [-1, -1, -1, -1, 3, -1, 5, 5]
    assert cases["ties"]["incl"] == [-1, 0, -1, 2, 3, 3, 5, 6]
    assert cases["intmax"]["strict"] == [-1, -1, 1, 1, -1, 4]
    assert cases["intmax"]["incl"] == [-1, -1, 1, 2, -1, 4]


# ------------------------------------------------------------ brute / verifier

@pytest.mark.parametrize("incl", [False, True])
@pytest.mark.parametrize("hi", [2, 16, 1000, 2**31 - 1])
def test_stack_equals_brute_force(incl, hi):
    rng = np.random.default_rng(1000 + hi)
    for trial in range(60):
        n = int(rng.integers(1, 700))
        a = rng.integers(0, hi, size=n, dtype=np.int64).astype(np.int32)
        if trial % 5 == 0:
            a[rng.random(n) < 0.3] = 2**31 - 1
        if trial % 7 == 0:
            a[rng.random(n) < 0.3] = -2**31
        exp = oracle.apsep_brute(a, incl)
        assert np.array_equal(oracle.apsep(a, incl), exp)
        assert oracle.verify(a, exp, incl) == -1


@pytest.mark.parametrize("incl", [False, True])
def test_verifier_rejects_planted_mistakes(incl):
    rng = np.random.default_rng(7)
    a = rng.integers(0, 50, size=400).astype(np.int32)
    good = oracle.apsep(a, incl)
    assert oracle.verify(a, good, incl) == -1
    idx = [i for i in range(1, 400) if good[i] >= 1][:20]
    for i in idx:
        for bad in (good[i] - 1, good[i] + 1, -1, i):
            b = good.copy()
            b[i] = bad
            if bad == good[i]:
                continue
            # a different index can only be right if it is also the answer
            assert oracle.verify(a, b, incl) == i, (i, bad)
    # strict vs incl confusion is caught on ties
    t = np.array([3, 3, 3, 1, 1], dtype=np.int32)
    assert oracle.verify(t, oracle.apsep(t, not incl), incl) != -1


def test_closed_forms_large():
    n = 1 << 20
    asc = oracle.fill("ascending", n)
    assert np.array_equal(oracle.apsep(asc), np.arange(-1, n - 1, dtype=np.int32))
    desc = oracle.fill("descending", n)
    assert (desc == n - np.arange(n)).all()
    assert (oracle.apsep(desc) == -1).all()
    saw = oracle.fill("sawtooth", n)
    i = np.arange(n)
    assert np.array_equal(oracle.apsep(saw), np.where(i % 4096 == 0, -1, i - 1))
    eq = oracle.fill("alleq", 4096)
    assert (oracle.apsep(eq) == -1).all()
    assert np.array_equal(oracle.apsep(eq, incl=True), np.arange(-1, 4095))


def test_blockdesc_structure():
    n = 1 << 18
    a = oracle.fill("blockdesc", n, seed=3)
    r = np.arange(n) >> 16
    # descending inside each run
    d = np.diff(a.astype(np.int64))
    inside = (r[1:] == r[:-1])
    assert (d[inside] == -1).all()
    out = oracle.apsep(a)
    assert oracle.verify(a, out) == -1


def test_random_classes_verify():
    for kind in ("random", "dup16", "dup2", "intmax"):
        a = oracle.fill(kind, 1 << 17, seed=11)
        out = oracle.apsep(a)
        assert oracle.verify(a, out) == -1, kind
        assert oracle.verify(a, oracle.apsep(a, True), True) == -1, kind
    r = oracle.fill("random", 1 << 16, seed=1)
    assert r.min() >= 0 and len(np.unique(r)) > 60000
    assert set(np.unique(oracle.fill("dup16", 4096, seed=2))) == set(range(16))
    im = oracle.fill("intmax", 1 << 16, seed=3)
    frac = (im == 2**31 - 1).mean()
    assert 0.2 < frac < 0.3


# ---------------------------------------------------------------- segmented

@pytest.mark.parametrize("incl", [False, True])
def test_segmented_equals_per_segment(incl):
    off = oracle.loguniform_offsets(5, 300, 1, 3000)
    assert off[0] == 0 and (np.diff(off) >= 1).all()
    a = oracle.fill("dup16", int(off[-1]), seed=9)
    got = oracle.apsep_segmented(a, off, incl)
    got_mt = oracle.apsep_segmented(a, off, incl, threads=4)
    assert np.array_equal(got, got_mt)
    for s in range(300):
        lo, hi = off[s], off[s + 1]
        assert np.array_equal(got[lo:hi], oracle.apsep(a[lo:hi], incl))
    # empty segments are allowed
    off2 = np.array([0, 0, 3, 3, 5], dtype=np.int64)
    a2 = np.array([5, 1, 2, 9, 8], dtype=np.int32)
    assert oracle.apsep_segmented(a2, off2).tolist() == [-1, -1, 1, -1, -1]


def test_loguniform_lengths():
    off = oracle.loguniform_offsets(1, 8192, 64, 2**20)
    L = np.diff(off)
    assert L.min() >= 64 and L.max() <= 2**20
    # log-uniform: median near sqrt(64 * 2^20) = 8192
    assert 4000 < np.median(L) < 16000


# ---------------------------------------------------------------- generators

def test_mt19937_standard_value():
    # [rand.predef]: the 10000th consecutive invocation of a default-
    # constructed std::mt19937 (seed 5489) produces 4123659995.
    r = oracle.MT19937(5489)
    for _ in range(9999):
        r()
    assert r() == 4123659995


def test_mt19937_fixture_and_stress_shape():
    with open(os.path.join(GOLDEN, "generators.json")) as f:
        g = json.load(f)
    for seed, vals in g["mt19937"].items():
        r = oracle.MT19937(int(seed))
        assert [r() for _ in range(8)] == vals
    # src/main.cu:146-152: n = 1 + rng() % maxN, then n values rng() % 1e6
    r = oracle.MT19937(42)
    first = r()
    n0 = 1 + first % 4096
    trials = oracle.stress_inputs(42, 3, 4096)
    assert len(trials[0]) == n0
    assert all(1 <= len(t) <= 4096 and t.min() >= 0 and t.max() < 1000000
               for t in trials)


def test_glibc_rand_matches_libc():
    libc = ctypes.CDLL("libc.so.6")
    with open(os.path.join(GOLDEN, "generators.json")) as f:
        g = json.load(f)
    for seed in (1, 5, 7, 99, 1234):
        libc.srand(seed)
        want = [libc.rand() for _ in range(5000)]
        got = oracle.glibc_rand(seed, 5000).tolist()
        assert got == want, seed
        assert got[:8] == g["glibc_rand"][str(seed)]
    # the classic first outputs of srand(1)
    assert oracle.glibc_rand(1, 2).tolist() == [1804289383, 846930886]


def test_glibc_intmax_matches_libc():
    """The INT_MAX-heavy input of src/bench_spt_le.cu:122-125, drawn here
    straight from libc's own rand(): the second draw happens only when the
    first is not a multiple of 4."""
    libc = ctypes.CDLL("libc.so.6")
    n = 20000
    libc.srand(5)
    want = []
    for _ in range(n):
        want.append(2147483647 if libc.rand() % 4 == 0 else libc.rand())
    got = oracle.glibc_intmax(5, n).tolist()
    assert got == want
    frac = sum(v == 2147483647 for v in got) / n
    assert 0.22 < frac < 0.28
