"""CPU oracle for the APSEP / ANSV-left hot path -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` import this package, and only as the checker.  The shipped
library (``apsep_amd``) never calls it.

The C restatement lives in ``oracle/apsep_oracle.c`` and is compiled to
``oracle/liboracle.so`` by :func:`build` (gcc -O3).  Every function there is
pinned by ``tests/test_oracle.py``:

* ``apsep`` (monotone stack, ``src/main.cu:102-113`` /
  ``src/bench_spt_le.cu:24-35``) against the reference's known-answer cases
  (``src/main.cu:181-222``), the O(n^2) definition scan and closed forms;
* ``verify`` (definition clauses) by planted mistakes;
* ``mt19937`` against the C++ standard's required 10000th output;
* ``glibc_rand`` against libc's own ``rand()``.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_SRC = os.path.join(_HERE, "apsep_oracle.c")
_LIB = os.path.join(_HERE, "liboracle.so")
_lib = None

# workload kinds understood by oracle_fill (see apsep_oracle.c)
KIND_RANDOM, KIND_DESC, KIND_ASC, KIND_DUP16, KIND_SAWTOOTH, KIND_BLOCKDESC, \
    KIND_ALLEQ, KIND_DUP2, KIND_INTMAX = range(9)
KIND_NAMES = {
    "random": KIND_RANDOM, "descending": KIND_DESC, "ascending": KIND_ASC,
    "dup16": KIND_DUP16, "sawtooth": KIND_SAWTOOTH,
    "blockdesc": KIND_BLOCKDESC, "alleq": KIND_ALLEQ, "dup2": KIND_DUP2,
    "intmax": KIND_INTMAX,
}


def build(force: bool = False) -> str:
    """Compile the C oracle into oracle/liboracle.so (gcc -O3)."""
    if force or not os.path.exists(_LIB) or \
            os.path.getmtime(_LIB) < os.path.getmtime(_SRC):
        tmp = _LIB + ".tmp%d" % os.getpid()
        cmd = ["gcc", "-O3", "-march=x86-64-v2", "-fPIC", "-shared",
               "-pthread", "-o", tmp, _SRC, "-lm"]
        subprocess.check_call(cmd)
        os.replace(tmp, _LIB)
    return _LIB


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        L = ctypes.CDLL(_LIB)
        i32p = ctypes.POINTER(ctypes.c_int32)
        i64p = ctypes.POINTER(ctypes.c_int64)
        L.oracle_apsep.argtypes = [i32p, ctypes.c_int64, ctypes.c_int, i32p]
        L.oracle_apsep_brute.argtypes = [i32p, ctypes.c_int64, ctypes.c_int, i32p]
        L.oracle_apsep_verify.argtypes = [i32p, ctypes.c_int64, ctypes.c_int, i32p]
        L.oracle_apsep_verify.restype = ctypes.c_int64
        L.oracle_apsep_segmented.argtypes = [i32p, i64p, ctypes.c_int32,
                                             ctypes.c_int, i32p]
        L.oracle_apsep_segmented_mt.argtypes = [i32p, i64p, ctypes.c_int32,
                                                ctypes.c_int, i32p,
                                                ctypes.c_int32]
        L.oracle_mt19937_state_size.restype = ctypes.c_int32
        L.oracle_mt19937_seed.argtypes = [ctypes.c_void_p, ctypes.c_uint32]
        L.oracle_mt19937_next.argtypes = [ctypes.c_void_p]
        L.oracle_mt19937_next.restype = ctypes.c_uint32
        L.oracle_stress_trial.argtypes = [ctypes.c_void_p, ctypes.c_uint32,
                                          i32p, ctypes.c_int32]
        L.oracle_stress_trial.restype = ctypes.c_int32
        L.oracle_glibc_rand_fill.argtypes = [ctypes.c_uint32, ctypes.c_int64, i32p]
        L.oracle_glibc_intmax_fill.argtypes = [ctypes.c_uint32, ctypes.c_int64, i32p]
        L.oracle_fill.argtypes = [ctypes.c_int32, ctypes.c_uint64,
                                  ctypes.c_int64, ctypes.c_int64,
                                  ctypes.c_int64, i32p]
        L.oracle_loguniform_offsets.argtypes = [ctypes.c_uint64, ctypes.c_int32,
                                                ctypes.c_double, ctypes.c_double,
                                                i64p]
        L.oracle_loguniform_offsets.restype = ctypes.c_int64
        _lib = L
    return _lib


def _p32(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))


def _p64(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int64))


def _as_i32(a):
    return np.ascontiguousarray(np.asarray(a), dtype=np.int32)


def apsep(a, incl: bool = False) -> np.ndarray:
    """Monotone-stack answers (int32, -1 = none)."""
    a = _as_i32(a)
    out = np.empty(a.shape[0], dtype=np.int32)
    if a.shape[0]:
        rc = lib().oracle_apsep(_p32(a), a.shape[0], int(incl), _p32(out))
        if rc:
            raise MemoryError("oracle_apsep: allocation failed")
    return out


def apsep_brute(a, incl: bool = False) -> np.ndarray:
    """O(n^2) scan from the definition (small n only)."""
    a = _as_i32(a)
    out = np.empty(a.shape[0], dtype=np.int32)
    lib().oracle_apsep_brute(_p32(a), a.shape[0], int(incl), _p32(out))
    return out


def verify(a, out, incl: bool = False) -> int:
    """-1 if ``out`` satisfies the definition for ``a``, else first bad index."""
    a = _as_i32(a)
    out = _as_i32(out)
    assert a.shape == out.shape
    r = lib().oracle_apsep_verify(_p32(a), a.shape[0], int(incl), _p32(out))
    if r == -2:
        raise MemoryError("oracle_apsep_verify: allocation failed")
    return int(r)


def apsep_segmented(a, offsets, incl: bool = False, threads: int = 1) -> np.ndarray:
    """Per-segment answers relative to each segment start."""
    a = _as_i32(a)
    off = np.ascontiguousarray(offsets, dtype=np.int64)
    nseg = off.shape[0] - 1
    out = np.empty(a.shape[0], dtype=np.int32)
    if threads <= 1:
        rc = lib().oracle_apsep_segmented(_p32(a), _p64(off), nseg, int(incl),
                                          _p32(out))
    else:
        rc = lib().oracle_apsep_segmented_mt(_p32(a), _p64(off), nseg,
                                             int(incl), _p32(out), threads)
    if rc:
        raise ValueError("oracle_apsep_segmented failed (rc=%d)" % rc)
    return out


class MT19937:
    """std::mt19937 restated in C (oracle_mt19937_*)."""

    def __init__(self, seed: int):
        self._buf = ctypes.create_string_buffer(lib().oracle_mt19937_state_size())
        lib().oracle_mt19937_seed(self._buf, seed & 0xFFFFFFFF)

    def __call__(self) -> int:
        return lib().oracle_mt19937_next(self._buf)

    def stress_trial(self, maxn: int) -> np.ndarray:
        """One trial of src/main.cu:146-152 (n = 1 + rng()%maxN, values
        rng()%1000000)."""
        cap = maxn
        vals = np.empty(cap, dtype=np.int32)
        n = lib().oracle_stress_trial(self._buf, maxn, _p32(vals), cap)
        return vals[:n].copy()


def stress_inputs(seed: int, trials: int, maxn: int):
    """The reference's seeded stress inputs (src/main.cu:144-175,246-252)."""
    rng = MT19937(seed)
    return [rng.stress_trial(maxn) for _ in range(trials)]


def glibc_rand(seed: int, n: int) -> np.ndarray:
    """``srand(seed)`` then ``n`` x ``rand()`` (glibc), restated."""
    out = np.empty(n, dtype=np.int32)
    lib().oracle_glibc_rand_fill(seed & 0xFFFFFFFF, n, _p32(out))
    return out


def glibc_intmax(seed: int, n: int) -> np.ndarray:
    """``srand(seed)``; ``h[i] = (rand() % 4 == 0) ? INT_MAX : rand()``
    (src/bench_spt_le.cu:122-125), restated."""
    out = np.empty(n, dtype=np.int32)
    if lib().oracle_glibc_intmax_fill(seed & 0xFFFFFFFF, n, _p32(out)):
        raise MemoryError("oracle_glibc_intmax_fill")
    return out


def fill(kind, n: int, seed: int = 0, base: int = 0, total: int | None = None,
         out: np.ndarray | None = None) -> np.ndarray:
    """Synthetic workload of one of the benchmark classes (see KIND_NAMES)."""
    if isinstance(kind, str):
        kind = KIND_NAMES[kind]
    if total is None:
        total = n
    if out is None:
        out = np.empty(n, dtype=np.int32)
    lib().oracle_fill(kind, seed, base, n, total, _p32(out))
    return out


def loguniform_offsets(seed: int, nseg: int, lo: float, hi: float) -> np.ndarray:
    off = np.empty(nseg + 1, dtype=np.int64)
    lib().oracle_loguniform_offsets(seed, nseg, lo, hi, _p64(off))
    return off
