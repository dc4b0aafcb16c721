"""Summarise rocprofv3 SQLite output (run_results.db) per kernel.

    python parallel-rdp_amd/rocprof_summary.py stats  <kernel-trace .db>
    python parallel-rdp_amd/rocprof_summary.py pmc    <pmc .db> [kernel-substring]

`stats` prints call count, total and average duration per kernel.  `pmc`
prints, per kernel and counter, the number of dispatches, the sum over
dispatches and the per-dispatch average (counter values are summed over the
per-SE/per-instance rows of one dispatch first).
"""
from __future__ import annotations

import collections
import sqlite3
import sys


def stats(path: str):
    db = sqlite3.connect(path)
    rows = db.execute("select name, total_calls, total_duration, average, percentage from top_kernels").fetchall()
    print(f"{'kernel':60s} {'calls':>7s} {'total_us':>12s} {'avg_us':>10s} {'pct':>6s}")
    for name, calls, total, avg, pct in rows:
        print(f"{name[:60]:60s} {calls:7d} {total / 1e3:12.1f} {avg / 1e3:10.3f} {pct:6.2f}")


def pmc(path: str, kernel_filter: str = ""):
    db = sqlite3.connect(path)
    per_dispatch = collections.defaultdict(float)
    names = {}
    for disp, kname, counter, value in db.execute(
            "select dispatch_id, kernel_name, counter_name, value from counters_collection"):
        if kernel_filter and kernel_filter not in kname:
            continue
        per_dispatch[(kname, counter, disp)] += value
        names[kname] = True
    agg = collections.defaultdict(lambda: [0, 0.0])
    for (kname, counter, disp), v in per_dispatch.items():
        a = agg[(kname, counter)]
        a[0] += 1
        a[1] += v
    print(f"{'kernel':48s} {'counter':22s} {'dispatches':>10s} {'sum':>16s} {'avg/dispatch':>16s}")
    for (kname, counter), (n, total) in sorted(agg.items()):
        short = kname.split("(")[0]
        print(f"{short[:48]:48s} {counter:22s} {n:10d} {total:16.1f} {total / n:16.1f}")


if __name__ == "__main__":
    if len(sys.argv) < 3:
        print(__doc__)
        sys.exit(1)
    if sys.argv[1] == "stats":
        stats(sys.argv[2])
    else:
        pmc(sys.argv[2], sys.argv[3] if len(sys.argv) > 3 else "")
