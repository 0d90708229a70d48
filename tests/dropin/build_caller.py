"""Build the reference-style C++ callers of the drop-in shims
(include/apsep/apsep.hpp), compiled by hipcc from tests/dropin/dropin_caller.cpp
and linked against apsep_amd/libapsep.so (run-path relative to the binary,
so the built pair travels together):
  tests/dropin/dropin_caller       the shims as shipped (the SPT names run
                                   the fastest solver)
  tests/dropin/dropin_caller_spt   APSEP_SHIM_SPT_ALGO=APSEP_ALGO_SPT: the SPT
                                   names run the single cooperative kernel"""
from __future__ import annotations

import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
SRC = os.path.join(HERE, "dropin_caller.cpp")
OUT = os.path.join(HERE, "dropin_caller")
OUT_SPT = os.path.join(HERE, "dropin_caller_spt")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")


def _build_one(out: str, defines: list) -> str:
    tmp = out + ".tmp%d" % os.getpid()
    cmd = [HIPCC, "-O2", "-std=c++17", "--offload-arch=gfx950", *defines,
           "-I", os.path.join(ROOT, "include"), SRC,
           "-L", os.path.join(ROOT, "apsep_amd"), "-lapsep",
           "-Wl,-rpath,$ORIGIN/../../apsep_amd", "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("hipcc (drop-in caller) failed (%d):\n%s" % (r.returncode, r.stderr))
    os.replace(tmp, out)
    return out


def build() -> list:
    return [_build_one(OUT, []),
            _build_one(OUT_SPT, ["-DAPSEP_SHIM_SPT_ALGO=APSEP_ALGO_SPT"])]


if __name__ == "__main__":
    print("\n".join(build()))
