"""Per-fragment VALU cost of each render mode (K_mode), measured on streams
whose every 8x8 tile is fully covered by every primitive (rdp_streams.
dense_stream), so no lane idles.

    rocprofv3 --pmc SQ_INSTS_VALU SQ_INSTS_SALU SQ_WAVES -d DIR -o run -- \\
        python3 parallel-rdp_amd/valu_calibration.py run OUT_RUN.json
    python3 parallel-rdp_amd/valu_calibration.py table DIR/run_results.db OUT_RUN.json OUT_TABLE.json

`run` renders each mode's stream once on the GPU (frame 0 clears, frame 1
draws the dense layers; one recorded replay, stages in launch order) and
writes the modes in dispatch order with their frame-1 fragment counts:
every layer covers every pixel once, so frame 1 has layers x width x height
fragments (tests/test_streams.py checks this count on the oracle).  `table` takes, per mode, the raster dispatch of frame 1 (the
last raster dispatch of the mode's replay; the clears before it are passes of
their own) and writes

    K_mode = SQ_INSTS_VALU x 64 / fragments      (lane-ops per fragment)

and the same for SALU (wave instructions per 64 fragments).  The bench
prices its workload with it: O = sum over modes of fragments x K_mode.
"""
from __future__ import annotations

import json
import os
import sqlite3
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

import numpy as np  # noqa: E402

import rdp_streams as rs  # noqa: E402


def run(out_path: str):
    import rdp_mi
    modes = []
    for mode in rs.DENSE_MODES:
        s = rs.dense_stream(mode)
        frags = s.num_triangles * s.width * s.height
        p = rdp_mi.CommandProcessor(rs.RDRAM_SIZE)
        p.hold_submissions()
        p.write_rdram(s.rdram_init if s.rdram_init is not None else np.zeros(rs.RDRAM_SIZE // 4, np.uint32))
        p.write_hidden_rdram(np.full(rs.RDRAM_SIZE // 2, s.hidden_init, np.uint8))
        for f in s.frames:
            p.enqueue_commands(f.words)
        rec = rdp_mi.Recording([p], batched=True)
        launches = rec.raster_launches
        rec.replay(1, use_graph=False)
        rec.close()
        p.close()
        modes.append({"mode": mode, "fragments": frags, "layers": s.num_triangles,
                      "framebuffer": f"{s.width}x{s.height}", "raster_launches": launches})
        print(f"[calibration] {mode}: {frags} fragments", flush=True)
    with open(out_path, "w") as fh:
        json.dump({"modes": modes}, fh, indent=1)


def _raster_dispatches(db_path: str, counter: str):
    """(kernel name, value) of every raster dispatch in dispatch order, the
    counter summed over its per-instance rows."""
    db = sqlite3.connect(db_path)
    q = ("select dispatch_id, kernel_name, sum(value) from counters_collection where counter_name = ? "
         "group by dispatch_id, kernel_name order by dispatch_id")
    return [(n, v) for _, n, v in db.execute(q, (counter,)).fetchall() if "k_raster" in n]


def table(db_path: str, run_path: str, out_path: str):
    with open(run_path) as fh:
        modes = json.load(fh)["modes"]
    valu = _raster_dispatches(db_path, "SQ_INSTS_VALU")
    salu = _raster_dispatches(db_path, "SQ_INSTS_SALU")
    expected = sum(m["raster_launches"] for m in modes)
    if len(valu) != expected or len(salu) != len(valu):
        raise RuntimeError(f"{len(valu)} raster dispatches, expected {expected}")
    out = {}
    last = -1
    for m in modes:
        last += m["raster_launches"]
        kname, v = valu[last]
        _, sv = salu[last]
        out[m["mode"]] = {"k_lane_ops_per_fragment": round(v * 64.0 / m["fragments"], 2),
                          "salu_per_64_fragments": round(sv * 64.0 / m["fragments"], 2),
                          "valu_wave_insts": int(v), "salu_wave_insts": int(sv), "fragments": m["fragments"],
                          "kernel": kname.split("(")[0], "stream": f"dense {m['mode']} {m['framebuffer']}, "
                                                                   f"{m['layers']} full-screen layers"}
    with open(out_path, "w") as fh:
        json.dump(out, fh, indent=1)
    for k, v in out.items():
        print(f"{k:12s} K = {v['k_lane_ops_per_fragment']:8.2f} lane-ops/fragment  "
              f"SALU {v['salu_per_64_fragments']:8.2f} per 64 fragments  ({v['kernel']})")


if __name__ == "__main__":
    if len(sys.argv) >= 3 and sys.argv[1] == "run":
        run(sys.argv[2])
    elif len(sys.argv) >= 5 and sys.argv[1] == "table":
        table(sys.argv[2], sys.argv[3], sys.argv[4])
    else:
        print(__doc__)
        sys.exit(1)
