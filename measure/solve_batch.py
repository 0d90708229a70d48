"""Repeated segmented solves of the batched configuration, for counter runs
of the segmented kernels.

  python measure/solve_batch.py [--nseg S] [--hi H] [--lanes L] [--reps R] [--copy C]
                                [--persistent] [--both]

The batch is bench.py's: S segments (default 8192) with log-uniform lengths
in [64, H] (H default 2^20), class per segment uniform over random /
descending / ascending, keys built on the device.  R segmented solves run
back to back on one plan with the tile shape fixed to L lanes (direct
launches): the four-kernel sequence, or with --persistent one persistent
launch per batch, or with --both R of each (sequence first).  Under
``rocprofv3 --pmc`` every dispatch of the segmented kernels then belongs to
this batch.  --copy C adds C plain device copies of the batch's keys (a
kernel of known traffic, 4 bytes per key each way) for scaling FETCH_SIZE /
WRITE_SIZE (measure/rocprof_summary.py traffic --n <keys>).  The key count
is printed.
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
    ap.add_argument("--nseg", type=int, default=8192)
    ap.add_argument("--lanes", type=int, default=256, choices=[128, 256, 512])
    ap.add_argument("--reps", type=int, default=5)
    ap.add_argument("--copy", type=int, default=0)
    ap.add_argument("--hi", type=int, default=1 << 20)
    ap.add_argument("--persistent", action="store_true")
    ap.add_argument("--both", action="store_true")
    a = ap.parse_args(argv)
    import torch

    import apsep_amd as A
    import bench
    x, off, _, _ = bench.batch_inputs(torch, a.nseg, 64, a.hi, 2024)
    total = int(off[-1])
    offs = torch.from_numpy(off).cuda()
    out = torch.empty(total, dtype=torch.int32, device="cuda")
    for _ in range(a.copy):
        out.copy_(x)
    with A.BatchPlan(total, a.nseg) as p:
        p.set_tile_lanes(a.lanes)
        p.use_graph(False)
        modes = (False, True) if a.both else (a.persistent,)
        for persist in modes:
            p.set_persistent(persist)
            for _ in range(a.reps):
                p.run(x, offs, out, total=total)
            torch.cuda.synchronize()
        p.check()
    print("solved batch of %d segments, %d keys, x%d, %d lanes, %s" % (
        a.nseg, total, a.reps, a.lanes,
        "both sequences" if a.both else "persistent" if a.persistent else "four kernels"),
        flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
