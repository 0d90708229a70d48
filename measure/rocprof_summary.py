"""Summaries of rocprofv3 result databases for the committed profiles.

  python measure/rocprof_summary.py kernels RESULTS.db [--command CMD]
      per-kernel dispatch count and duration (avg / min / max, us) from a
      ``--kernel-trace`` run.

  python measure/rocprof_summary.py traffic FETCH.db WRITE.db --n N
          [--algo wstl] [--json OUT] [--command CMD]
      HBM bytes per launch of each solver kernel from two ``--pmc`` runs
      (FETCH_SIZE, then WRITE_SIZE; KiB per dispatch).  Both counters are
      scaled by what they read for a plain device copy of N int32 in the
      same run (the copy moves exactly 4N bytes each way), so the result is
      in true bytes.

  python measure/rocprof_summary.py counters DB [DB ...] --kernel NAME
          [--skip K] [--command CMD]
      mean per dispatch of every counter recorded in the given ``--pmc``
      databases for the kernels whose name contains NAME (the first K
      dispatches skipped), and per wave when SQ_WAVES is among them.
"""
from __future__ import annotations

import argparse
import json
import sqlite3
import statistics
import sys


def _short(name: str) -> str:
    name = name.replace("void ", "")
    return name.split("(")[0][:70]


def kernel_stats(db: str):
    con = sqlite3.connect(db)
    rows = con.execute(
        "select name, count(*), avg(end - start), min(end - start), max(end - start) "
        "from kernels group by name order by sum(end - start) desc").fetchall()
    return [(_short(r[0]), r[1], r[2] / 1e3, r[3] / 1e3, r[4] / 1e3) for r in rows]


def pmc_per_dispatch(db: str, counter: str):
    """{kernel short name: [value per dispatch]} for one counter."""
    con = sqlite3.connect(db)
    rows = con.execute(
        "select dispatch_id, name, sum(counter_value) from pmc_events "
        "where counter_name = ? group by dispatch_id", (counter,)).fetchall()
    out: dict[str, list[float]] = {}
    for _, name, v in rows:
        out.setdefault(_short(name), []).append(float(v))
    return out


def counter_names(db: str) -> list[str]:
    con = sqlite3.connect(db)
    return [r[0] for r in con.execute("select distinct counter_name from pmc_events")]


def _copy_kernel(per: dict[str, list[float]]):
    """The calibration copy: the copy kernel with the most dispatches."""
    cands = [(len(v), k) for k, v in per.items() if "copy" in k.lower()]
    return max(cands)[1] if cands else None


def main(argv=None):
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    k = sub.add_parser("kernels")
    k.add_argument("db")
    k.add_argument("--command", default="")
    t = sub.add_parser("traffic")
    t.add_argument("fetch_db")
    t.add_argument("write_db")
    t.add_argument("--n", type=int, required=True)
    t.add_argument("--algo", default="wstl")
    t.add_argument("--json", default=None)
    t.add_argument("--command", default="")
    c = sub.add_parser("counters")
    c.add_argument("dbs", nargs="+")
    c.add_argument("--kernel", required=True)
    c.add_argument("--skip", type=int, default=2)
    c.add_argument("--command", default="")
    a = ap.parse_args(argv)

    if a.cmd == "counters":
        vals = {}
        for db in a.dbs:
            for cn in counter_names(db):
                for name, v in pmc_per_dispatch(db, cn).items():
                    if a.kernel in name:
                        v = v[a.skip:] if len(v) > a.skip else v
                        vals[cn] = statistics.mean(v)
        if a.command:
            print("command: %s" % a.command)
        waves = vals.get("SQ_WAVES")
        print("%-24s %16s %12s" % ("counter", "per dispatch", "per wave"))
        for cn in sorted(vals):
            pw = "%12.1f" % (vals[cn] / waves) if waves else ""
            print("%-24s %16.0f %s" % (cn, vals[cn], pw))
        return 0

    if a.cmd == "kernels":
        if a.command:
            print("command: %s" % a.command)
        print("%-70s %6s %10s %10s %10s" % ("kernel", "calls", "avg_us", "min_us", "max_us"))
        for name, cnt, avg, mn, mx in kernel_stats(a.db):
            print("%-70s %6d %10.2f %10.2f %10.2f" % (name, cnt, avg, mn, mx))
        return 0

    fetch = pmc_per_dispatch(a.fetch_db, "FETCH_SIZE")
    write = pmc_per_dispatch(a.write_db, "WRITE_SIZE")
    copy_f, copy_w = _copy_kernel(fetch), _copy_kernel(write)
    if copy_f is None or copy_w is None:
        print("no copy kernel found for calibration", file=sys.stderr)
        return 1
    true_bytes = 4.0 * a.n
    f_scale = statistics.median(fetch[copy_f]) * 1024.0 / true_bytes
    w_scale = statistics.median(write[copy_w]) * 1024.0 / true_bytes
    if a.command:
        print("command: %s" % a.command)
    print("calibration: %s, %d int32 copied; FETCH_SIZE reads %.4f x and "
          "WRITE_SIZE %.4f x the true bytes" % (copy_f, a.n, f_scale, w_scale))
    print("%-50s %6s %14s %14s %14s %10s" % ("kernel", "calls", "fetch_B", "write_B",
                                             "total_B", "B/elem"))
    res = {"n": a.n, "algo": a.algo, "fetch_scale": f_scale, "write_scale": w_scale,
           "kernels": {}}
    for name in sorted(set(fetch) & set(write)):
        if "apsep" not in name:
            continue
        fb = statistics.mean(fetch[name]) * 1024.0 / f_scale
        wb = statistics.mean(write[name]) * 1024.0 / w_scale
        res["kernels"][name] = {"calls": len(fetch[name]), "fetch_bytes": fb,
                                "write_bytes": wb}
        print("%-50s %6d %14.0f %14.0f %14.0f %10.3f" % (name, len(fetch[name]), fb, wb,
                                                        fb + wb, (fb + wb) / a.n))
    tile = [v for k, v in res["kernels"].items() if "k_tile_pass" in k]
    if tile:
        res["tile_pass_bytes_per_launch"] = round(tile[0]["fetch_bytes"] + tile[0]["write_bytes"])
    if a.json:
        with open(a.json, "w") as f:
            json.dump(res, f, indent=1)
    return 0


if __name__ == "__main__":
    sys.exit(main())
