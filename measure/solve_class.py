"""Repeated solves of one input class, for profiling a single class.

  python measure/solve_class.py KIND [--n N] [--reps R] [--algo ALGO] [--copy C]

KIND is random (uniform keys in [0, 2^31 - 1), device generator seeded
2024), descending (a[i] = N - i) or ascending (a[i] = i).  The keys are
generated on the device, then R solves run back to back on one scratch.
Under ``rocprofv3 --pmc`` every dispatch of the solver's kernels then
belongs to that class.  --copy C adds C plain device copies of N int32 (a
kernel of known traffic, 4N bytes each way) for scaling FETCH_SIZE /
WRITE_SIZE (measure/rocprof_summary.py traffic).
"""
from __future__ import annotations

import argparse
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("kind", choices=["random", "descending", "ascending"])
    ap.add_argument("--n", type=int, default=1 << 25)
    ap.add_argument("--reps", type=int, default=10)
    ap.add_argument("--algo", default="wstl")
    ap.add_argument("--copy", type=int, default=0)
    a = ap.parse_args(argv)
    import torch

    import apsep_amd as A
    n = a.n
    if a.kind == "random":
        g = torch.Generator(device="cuda")
        g.manual_seed(2024)
        x = torch.randint(0, 2**31 - 1, (n,), dtype=torch.int32, device="cuda", generator=g)
    elif a.kind == "descending":
        x = n - torch.arange(n, dtype=torch.int32, device="cuda")
    else:
        x = torch.arange(n, dtype=torch.int32, device="cuda")
    out = torch.empty(n, dtype=torch.int32, device="cuda")
    for _ in range(a.copy):
        out.copy_(x)
    with A.Scratch(a.algo, n) as s:
        for _ in range(a.reps):
            s.run(x, out)
        torch.cuda.synchronize()
    print("solved %s n=%d x%d (%s)" % (a.kind, n, a.reps, a.algo), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
