"""Summarise rocprofv3 SQLite output (<name>_results.db) per kernel.

    python parallel-rdp_amd/rocprof_summary.py stats  <kernel-trace .db> [kernel-substring]
    python parallel-rdp_amd/rocprof_summary.py pmc    <pmc .db> [kernel-substring]
    python parallel-rdp_amd/rocprof_summary.py gaps   <kernel-trace .db> <step-marker kernel> <marker grid_x>
    python parallel-rdp_amd/rocprof_summary.py traffic <fetch .db> <write .db> <copy-fetch .db> <copy-write .db>
                                                       <sq .db> <copy bytes> <out .json> [kernel-substring]

`traffic` writes, per kernel matching the substring (default k_raster), the
HBM bytes per dispatch (FETCH_SIZE and WRITE_SIZE in KiB, each scaled by the
factor that makes the same counter read the known byte count of a plain copy
kernel: the copy runs' largest kernel) and the SQ instruction counts per
dispatch, as JSON keyed by the kernel's short name (k_raster<0>, ...).

`gaps` splits the trace into steps at each launch of the marker kernel with
the given grid size (one per step) and prints, over the steps, the median of
the step's wall time, of the time at least one kernel runs (the union of the
kernels' intervals) and of the idle time between kernels: the most a single
persistent launch per step could take off the step.

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


def gaps(path: str, marker: str, marker_grid: int):
    db = sqlite3.connect(path)
    rows = sorted(db.execute("select start, end, name, grid_x from kernels").fetchall())
    starts = [s for s, e, n, g in rows if marker in n and g == marker_grid]
    steps = []
    for a, b in zip(starts, starts[1:]):
        iv = [(s, e) for s, e, n, g in rows if a <= s < b]
        if not iv:
            continue
        end = max(e for s, e in iv)
        busy, cur_s, cur_e = 0, None, None
        for s, e in iv:
            if cur_e is None or s > cur_e:
                if cur_e is not None:
                    busy += cur_e - cur_s
                cur_s, cur_e = s, e
            else:
                cur_e = max(cur_e, e)
        busy += cur_e - cur_s
        steps.append((end - a, busy, len(iv)))
    if not steps:
        print("no steps found")
        return
    med = lambda xs: sorted(xs)[len(xs) // 2]
    wall = med([w for w, b, k in steps])
    busy = med([b for w, b, k in steps])
    idle = med([w - b for w, b, k in steps])
    kernels = med([k for w, b, k in steps])
    print(f"steps {len(steps)}  kernels per step {kernels}")
    print(f"median step wall {wall / 1e3:.1f} us, kernels running {busy / 1e3:.1f} us, "
          f"idle between kernels {idle / 1e3:.1f} us ({100.0 * idle / wall:.1f} %)")


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


def _per_dispatch(path: str, counter: str, kernel_filter: str = ""):
    """{short kernel name: average of `counter` per dispatch}."""
    db = sqlite3.connect(path)
    per = collections.defaultdict(float)
    for disp, kname, cname, value in db.execute(
            "select dispatch_id, kernel_name, counter_name, value from counters_collection"):
        if cname == counter and (not kernel_filter or kernel_filter in kname):
            per[(short_name(kname), disp)] += value
    agg = collections.defaultdict(list)
    for (k, _), v in per.items():
        agg[k].append(v)
    return {k: sum(v) / len(v) for k, v in agg.items()}


def short_name(kname: str) -> str:
    """'void rdp_mi::k_raster<0u, false, false>(...)' -> 'k_raster<0>' (8x8
    tiles); 'void rdp_mi::k_raster<0u, true, false>(...)' -> 'k_raster<0>
    16x16'; 'void rdp_mi::k_raster<0u, false, true>(...)' -> 'k_raster<0>
    upscaled'.  Names without the third argument (older builds) map alike."""
    n = kname.split("(")[0].replace("void ", "").replace("rdp_mi::", "")
    if "<" not in n:
        return n
    base, args = n.split("<", 1)
    parts = [a.strip() for a in args.rstrip(">").split(",")]
    quad = len(parts) > 1 and parts[1] == "true"
    up = len(parts) > 2 and parts[2] == "true"
    n = f"{base}<{parts[0].rstrip('u')}>"
    return n + (" 16x16" if quad else "") + (" upscaled" if up else "")


def traffic(fetch_db, write_db, copyf_db, copyw_db, sq_db, copy_bytes, out, kernel_filter="k_raster"):
    import json
    fetch = _per_dispatch(fetch_db, "FETCH_SIZE", kernel_filter)
    write = _per_dispatch(write_db, "WRITE_SIZE", kernel_filter)
    cf = max(_per_dispatch(copyf_db, "FETCH_SIZE").values())
    cw = max(_per_dispatch(copyw_db, "WRITE_SIZE").values())
    ff, wf = copy_bytes / (cf * 1024), copy_bytes / (cw * 1024)
    res = {}
    for k in sorted(fetch):
        res[k] = {"fetch_kib_per_dispatch": round(fetch[k], 1), "fetch_factor": round(ff, 3),
                  "write_kib_per_dispatch": round(write.get(k, 0.0), 1), "write_factor": round(wf, 3),
                  "traffic_bytes_per_launch": int(fetch[k] * 1024 * ff + write.get(k, 0.0) * 1024 * wf)}
    for counter, key in (("SQ_INSTS_VALU", "valu_insts_per_launch"), ("SQ_INSTS_SALU", "salu_insts_per_launch"),
                         ("SQ_INSTS_LDS", "lds_insts_per_launch"), ("SQ_LDS_BANK_CONFLICT", "lds_bank_conflict_cycles")):
        for k, v in _per_dispatch(sq_db, counter, kernel_filter).items():
            res.setdefault(k, {})[key] = int(v)
    with open(out, "w") as fh:
        json.dump(res, fh, indent=1)
    print(json.dumps(res, indent=1))


if __name__ == "__main__":
    if len(sys.argv) > 1 and sys.argv[1] == "traffic":
        a = sys.argv[2:]
        traffic(a[0], a[1], a[2], a[3], a[4], float(a[5]), a[6], *(a[7:8]))
        sys.exit(0)
    if len(sys.argv) < 3:
        print(__doc__)
        sys.exit(1)
    flt = sys.argv[3] if len(sys.argv) > 3 else ""
    if sys.argv[1] == "gaps":
        gaps(sys.argv[2], sys.argv[3], int(sys.argv[4]))
    elif sys.argv[1] == "stats":
        stats(sys.argv[2], flt)
    else:
        pmc(sys.argv[2], flt)
