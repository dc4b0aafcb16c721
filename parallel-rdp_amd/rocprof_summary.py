"""Summarise rocprofv3 SQLite output (<name>_results.db) per kernel.

    python parallel-rdp_amd/rocprof_summary.py stats  <kernel-trace .db> [kernel-substring]
    python parallel-rdp_amd/rocprof_summary.py pmc    <pmc .db> [kernel-substring]

`stats` prints, per kernel, the call count and total/average duration, then
for kernels matching the substring a breakdown by launch grid size and the
averages over launches that ran alone (no other kernel overlapping in time),
which are the ones comparable with per-stage event timing.  `pmc`
prints, per kernel and counter, the number of dispatches, the sum over
dispatches and the per-dispatch average (counter values are first summed over
the per-instance rows of one dispatch).  FETCH_SIZE and WRITE_SIZE are in KiB.
"""
from __future__ import annotations

import collections
import sqlite3
import sys


def stats(path: str, kernel_filter: str = ""):
    db = sqlite3.connect(path)
    rows = db.execute("select name, start, end, grid_x from kernels").fetchall()
    per = collections.defaultdict(list)
    grids = collections.defaultdict(list)
    for name, start, end, grid in rows:
        short = name.split("(")[0]
        per[short].append(end - start)
        if kernel_filter and kernel_filter in name:
            grids[(short, grid)].append(end - start)
    total_all = sum(sum(v) for v in per.values())
    print(f"{'kernel':48s} {'calls':>7s} {'total_ms':>10s} {'avg_us':>10s} {'pct':>6s}")
    for short, d in sorted(per.items(), key=lambda kv: -sum(kv[1])):
        print(f"{short[:48]:48s} {len(d):7d} {sum(d) / 1e6:10.3f} {sum(d) / len(d) / 1e3:10.2f} "
              f"{100.0 * sum(d) / total_all:6.2f}")
    if kernel_filter:
        # Launches that overlap no other kernel in time.
        iv = sorted((start, end, name.split("(")[0]) for name, start, end, _ in rows)
        alone = collections.defaultdict(list)
        reach = -1
        for i, (start, end, short) in enumerate(iv):
            nxt = iv[i + 1][0] if i + 1 < len(iv) else None
            if start >= reach and (nxt is None or nxt >= end) and kernel_filter in short:
                alone[short].append(end - start)
            reach = max(reach, end)
        if alone:
            print(f"\n{'kernel (launches running alone)':48s} {'calls':>7s} {'avg_us':>10s}")
            for short, d in sorted(alone.items()):
                print(f"{short[:48]:48s} {len(d):7d} {sum(d) / len(d) / 1e3:10.2f}")
            allv = [x for d in alone.values() for x in d]
            print(f"{'all matching kernels':48s} {len(allv):7d} {sum(allv) / len(allv) / 1e3:10.2f}")
            grid_of = {(start, end): g for name, start, end, g in rows}
            by_grid = collections.defaultdict(list)
            reach = -1
            for i, (start, end, short) in enumerate(iv):
                nxt = iv[i + 1][0] if i + 1 < len(iv) else None
                if start >= reach and (nxt is None or nxt >= end) and kernel_filter in short:
                    by_grid[(short, grid_of[(start, end)])].append(end - start)
                reach = max(reach, end)
            print(f"\n{'kernel (alone)':32s} {'grid_x (threads)':>16s} {'calls':>7s} {'avg_us':>10s}")
            for (short, grid), d in sorted(by_grid.items(), key=lambda kv: (kv[0][0], -kv[0][1])):
                print(f"{short[:32]:32s} {grid:16d} {len(d):7d} {sum(d) / len(d) / 1e3:10.2f}")
    if grids:
        print(f"\n{'kernel':32s} {'grid_x (threads)':>16s} {'calls':>7s} {'avg_us':>10s} {'min_us':>10s} {'max_us':>10s}")
        for (short, grid), d in sorted(grids.items(), key=lambda kv: (kv[0][0], -kv[0][1])):
            print(f"{short[:32]:32s} {grid:16d} {len(d):7d} {sum(d) / len(d) / 1e3:10.2f} {min(d) / 1e3:10.2f} "
                  f"{max(d) / 1e3:10.2f}")


def pmc(path: str, kernel_filter: str = ""):
    db = sqlite3.connect(path)
    per_dispatch = collections.defaultdict(float)
    for disp, kname, counter, value in db.execute(
            "select dispatch_id, kernel_name, counter_name, value from counters_collection"):
        if kernel_filter and kernel_filter not in kname:
            continue
        per_dispatch[(kname, counter, disp)] += value
    agg = collections.defaultdict(lambda: [0, 0.0])
    for (kname, counter, disp), v in per_dispatch.items():
        a = agg[(kname, counter)]
        a[0] += 1
        a[1] += v
    print(f"{'kernel':48s} {'counter':14s} {'dispatches':>10s} {'sum':>16s} {'avg/dispatch':>16s}")
    for (kname, counter), (n, total) in sorted(agg.items()):
        short = kname.split("(")[0]
        print(f"{short[:48]:48s} {counter:14s} {n:10d} {total:16.1f} {total / n:16.1f}")


if __name__ == "__main__":
    if len(sys.argv) < 3:
        print(__doc__)
        sys.exit(1)
    flt = sys.argv[3] if len(sys.argv) > 3 else ""
    if sys.argv[1] == "stats":
        stats(sys.argv[2], flt)
    else:
        pmc(sys.argv[2], flt)
