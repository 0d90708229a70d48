"""Build tests/dropin/dropin_caller: the reference-style C++ caller of the
drop-in shims (include/apsep/apsep.hpp), compiled by hipcc and linked
against apsep_amd/libapsep.so (run-path relative to the binary, so the
built pair travels together)."""
from __future__ import annotations

import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
SRC = os.path.join(HERE, "dropin_caller.cpp")
OUT = os.path.join(HERE, "dropin_caller")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")


def build() -> str:
    tmp = OUT + ".tmp%d" % os.getpid()
    cmd = [HIPCC, "-O2", "-std=c++17", "--offload-arch=gfx950",
           "-I", os.path.join(ROOT, "include"), SRC,
           "-L", os.path.join(ROOT, "apsep_amd"), "-lapsep",
           "-Wl,-rpath,$ORIGIN/../../apsep_amd", "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("hipcc (drop-in caller) failed (%d):\n%s" % (r.returncode, r.stderr))
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build())
