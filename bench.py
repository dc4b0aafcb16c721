#!/usr/bin/env python3
"""Benchmark of the RDP rendering hot path on one MI355X.

Workload (the batched form): 64 independent seeded command streams, each with
its own 8 MB device RDRAM; triangle counts log-uniform in [32, 8192];
framebuffers 320x240 or 640x480 at 16 or 32 bits; render modes drawn from the
plain shaded, textured and 2-cycle (CI textures, mip LOD, fog, coverage AA)
workloads.  One step renders one frame of every
stream.

The stream commands are decoded by the host front end and uploaded once
(rdp_mi_record); the timed region replays the recorded passes from device
memory as one captured hipGraph per step.  The plain path (one launch per
stage per render pass, one stream at a time, no graph) is timed the same way
for comparison.  After the timed loop every stream's RDRAM, hidden RDRAM and
TMEM is compared byte for byte with the CPU oracle; only bit-exact frames
count toward frames/s.

Serving form (reported under "serving"): 256 instances with device-resident
state send frame requests to the continuous-batching server (host decode on
a thread pool, batched submission, admission by a device-byte budget) as a
Poisson process at 25/50/90% of the measured peak; p50/p99 latency is
completion minus arrival, and every instance's final state is checked
against the oracle.

Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "parallel-rdp_amd"))

import rdp_streams as rs  # noqa: E402

HBM_PEAK_GBS = 8000.0
VALU_PEAK_TOPS = 39.3  # 256 CUs x 4 SIMDs x 16 lanes x 2.4 GHz, 32-bit integer lane operations


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def setup_processors(rdp_mi, streams, flags=0):
    procs = []
    for s in streams:
        p = rdp_mi.CommandProcessor(rs.RDRAM_SIZE, flags=flags)
        # Every enqueued pass stays pending for the recording (no idle-time
        # submission behind the recording's back).
        p.hold_submissions()
        p.write_rdram(s.rdram_init if s.rdram_init is not None else np.zeros(rs.RDRAM_SIZE // 4, np.uint32))
        p.write_hidden_rdram(np.full(rs.RDRAM_SIZE // 2, s.hidden_init, np.uint8))
        procs.append(p)
    return procs


def enqueue_all(procs, streams):
    for p, s in zip(procs, streams):
        for f in s.frames:
            p.enqueue_commands(f.words)


def _oracle_module():
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    return oracle


def oracle_stream(s, digests=False, scanout_frames=False):
    """CPU oracle over one stream: final (RDRAM, hidden RDRAM, TMEM), the
    state digest after every frame (digests) and the VI image of every frame
    that carries VI registers (scanout_frames), the seconds it took and its
    fragments by render mode."""
    oracle = _oracle_module()
    t0 = time.perf_counter()
    o = oracle.Oracle(rs.RDRAM_SIZE)
    if s.rdram_init is not None:
        o.rdram[:] = s.rdram_init
    o.hidden[:] = s.hidden_init
    per_frame, images = [], []
    for f in s.frames:
        o.process(f.words)
        if scanout_frames and f.vi:
            for r, v in f.vi.items():
                o.set_vi_register(r, v)
            images.append(o.scanout())
        if digests:
            o.flush()
            per_frame.append(o.digest())
    o.flush()
    out = (o.rdram.copy(), o.hidden.copy(), o.tmem.copy())
    elapsed = time.perf_counter() - t0
    frags = o.fragments_by_mode()
    o.close()
    return out, per_frame, images, elapsed, frags


def cpu_threads():
    """Host threads the CPU baseline may use: this process's CPU affinity."""
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


def cpu_model():
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def oracle_results(streams, threads):
    """CPU oracle over every batch stream on `threads` host threads (the
    ctypes calls release the GIL; each stream is independent), which is the
    all-core CPU baseline and the reference the GPU results are checked
    against; then a one-core sample: the first 16 streams in sequence."""
    from concurrent.futures import ThreadPoolExecutor
    t0 = time.perf_counter()
    with ThreadPoolExecutor(max_workers=threads) as ex:
        outs = list(ex.map(oracle_stream, streams))
    wall = time.perf_counter() - t0
    sample = streams[:16]
    t0 = time.perf_counter()
    for s in sample:
        oracle_stream(s)
    one_core = time.perf_counter() - t0
    frags = {}
    for o in outs:
        for k, v in o[4].items():
            frags[k] = frags.get(k, 0) + v
    return [o[0] for o in outs], wall, len(sample), one_core, frags


def exact_frames(procs, refs):
    exact = 0
    mismatched_bytes = 0
    for p, r in zip(procs, refs):
        n = int(np.count_nonzero(p.read_rdram().view(np.uint8) != r[0].view(np.uint8)))
        n += int(np.count_nonzero(p.read_hidden_rdram() != r[1]))
        n += int(np.count_nonzero(p.read_tmem() != r[2]))
        mismatched_bytes += n
        exact += n == 0
    return exact, mismatched_bytes


def serving(rdp_mi, streams, refs, args):
    """Serving form: `instances` emulator instances, each with its own
    processor (8 MB RDRAM, hidden RDRAM and TMEM on the device); instance i
    renders the frame of batch stream i mod len(streams) over and over.
    Requests arrive as a Poisson process at 25/50/90% of the peak rate the
    server sustains with every request queued at once; latency is completion
    minus arrival.  Every batch stream's frame is idempotent (rendering it
    again leaves RDRAM, hidden RDRAM and TMEM unchanged), so after all runs
    each instance must hold exactly the oracle result of its stream."""
    rng = np.random.default_rng(0x5E5E)
    n = args.instances
    tmpl = [i % len(streams) for i in range(n)]
    procs = setup_processors(rdp_mi, [streams[t] for t in tmpl])
    words = [s.frames[0].words for s in streams]
    budget = args.hbm_budget_mib << 20

    # One server for every run: after the first (peak) run its threads,
    # streams and arenas are warm, and each later run starts on a reset clock.
    srv = rdp_mi.Server(procs, hbm_budget_bytes=budget, max_batch=args.max_batch, streams=args.serve_streams,
                        decode_threads=args.serve_decode_threads)
    srv.start()
    lead_us = 300000.0  # arrivals count from 300 ms after the reset, after all submits

    def run(arrivals, instances, reset=True):
        if reset:
            srv.reset(lead_us)
        for a, i in zip(arrivals, instances):
            srv.submit(int(i), words[tmpl[int(i)]], arrival_us=float(a))
        srv.drain()
        return srv.results(), srv.stats()

    # Peak: every request queued at once (a warm-up pass first).
    count = n * args.peak_frames_per_instance
    run(np.zeros(n), np.arange(n))
    res, st = run(np.zeros(count), np.tile(np.arange(n), args.peak_frames_per_instance))
    peak = count / ((res["complete_us"].max() - res["arrival_us"].min()) * 1e-6)
    log(f"[serve] peak {peak:.0f} frames/s over {count} frames, {st['batches']} batches "
        f"(largest {st['max_batch_frames']} frames, {st['max_arena_bytes'] / 2**20:.1f} MiB arena)")
    out = {"instances": n, "distinct_streams": len(streams), "hbm_budget_bytes": budget,
           "submission_streams": args.serve_streams, "max_batch": args.max_batch,
           "decode_threads": args.serve_decode_threads,
           "resident_bytes_per_instance": rs.RDRAM_SIZE + rs.RDRAM_SIZE // 2 + 4096,
           "peak_frames_per_s": round(peak, 1), "peak_stats": st, "loads": []}
    for frac in (0.25, 0.5, 0.9):
        rate = frac * peak
        m = int(min(args.serve_max_frames, max(256, rate * args.serve_seconds)))
        arrivals = np.cumsum(rng.exponential(1e6 / rate, size=m))
        inst = rng.integers(0, n, size=m)
        res, st = run(arrivals, inst)
        lat = (res["complete_us"] - res["arrival_us"]) / 1e3
        span = (res["complete_us"].max() - res["arrival_us"].min()) * 1e-6
        entry = {"offered_frac": frac, "offered_frames_per_s": round(rate, 1), "frames": m,
                 "achieved_frames_per_s": round(m / span, 1),
                 "p50_ms": round(float(np.percentile(lat, 50)), 4), "p99_ms": round(float(np.percentile(lat, 99)), 4),
                 "max_ms": round(float(lat.max()), 4), "mean_batch_frames": round(st["frames"] / max(st["batches"], 1), 2),
                 "max_arena_bytes": st["max_arena_bytes"],
                 "submit_ms_mean": round(st["submit_us_total"] / max(st["batches"], 1) / 1e3, 4),
                 "submit_ms_max": round(st["submit_us_max"] / 1e3, 4)}
        for key, d in (("queue", res["admit_us"] - res["arrival_us"]), ("decode", res["submit_us"] - res["admit_us"]),
                       ("gpu", res["complete_us"] - res["submit_us"])):
            entry[f"{key}_p50_ms"] = round(float(np.percentile(d, 50)) / 1e3, 4)
            entry[f"{key}_p99_ms"] = round(float(np.percentile(d, 99)) / 1e3, 4)
        if args.serve_dump:
            np.savez(f"{args.serve_dump}_{int(frac * 100)}.npz", words=np.array([words[tmpl[int(i)]].size
                                                                                 for i in res["instance"]]), **res)
        log(f"[serve] {frac:.0%} load: {entry}")
        out["loads"].append(entry)
    srv.close()
    exact, mism = exact_frames(procs, [refs[t] for t in tmpl])
    out["bit_exact_instances"] = f"{exact}/{n}"
    out["mismatched_bytes"] = mism
    log(f"[serve] final state bit-exact {exact}/{n} ({mism} mismatched bytes)")
    for p in procs:
        p.close()
    return out


def oracle_upscaled(s, factor, ssaa):
    """The oracle's upscaled run of one stream: native RDRAM, hidden RDRAM,
    TMEM, the RDRAM and hidden sample planes."""
    o = _oracle_module().Oracle(rs.RDRAM_SIZE)
    o.enable_upscaling(factor)
    if ssaa is not None:
        o.enable_super_sampled_readback(ssaa)
    if s.rdram_init is not None:
        o.rdram[:] = s.rdram_init
    o.hidden[:] = s.hidden_init
    for f in s.frames:
        o.process(f.words)
    o.flush()
    out = (o.rdram.copy(), o.hidden.copy(), o.tmem.copy(), o.upscaled_rdram.copy(), o.upscaled_hidden.copy())
    o.close()
    return out


def upscaled_rows(rdp_mi, streams, args):
    """The batch rendered with upscaling (2x, 4x) and with super-sampled
    readback (2x, dithered), replayed as one graph per step like the
    headline form.  A sample of the streams is checked against the oracle's
    upscaled run, native and sample-plane memory byte for byte; frames/s
    counts the step's frames only when every sampled stream is exact."""
    rows = {}
    sample = [i for i in (0, 17, 33, 50) if i < len(streams)]
    for name, factor, ssaa in (("2x", 2, None), ("4x", 4, None), ("2x_ssaa_dither", 2, True)):
        flags = rdp_mi.FLAG_UPSCALING[factor]
        if ssaa is not None:
            flags |= rdp_mi.FLAG_SUPER_SAMPLED_READ_BACK | (rdp_mi.FLAG_SUPER_SAMPLED_DITHER if ssaa else 0)
        procs = setup_processors(rdp_mi, streams, flags)
        enqueue_all(procs, streams)
        rec = rdp_mi.Recording(procs, batched=True)
        rec.replay(max(args.warmup, 1), use_graph=True)
        ms = float(rec.replay(max(3, args.steps // 5), use_graph=True).mean())
        exact = 0
        for i in sample:
            p = procs[i]
            gpu = (p.read_rdram(), p.read_hidden_rdram(), p.read_tmem(), p.read_upscaled_rdram(),
                   p.read_upscaled_hidden_rdram())
            ref = oracle_upscaled(streams[i], factor, ssaa)
            exact += all(np.array_equal(np.ascontiguousarray(a).view(np.uint8), np.ascontiguousarray(b).view(np.uint8))
                         for a, b in zip(gpu, ref))
        rows[name] = {"factor": factor, "super_sampled_readback": ssaa is not None, "dither": bool(ssaa),
                      "ms_per_step": round(ms, 4),
                      "frames_per_s": round(len(streams) / (ms / 1e3), 1) if exact == len(sample) else 0.0,
                      "bit_exact_sample": f"{exact}/{len(sample)}", "passes": rec.passes, "rounds": rec.rounds}
        log(f"[upscaled] {name}: {rows[name]}")
        rec.close()
        for p in procs:
            p.close()
    return rows


def config_rows(rdp_mi, pending, args):
    """Configs 0-2 on their own: each one stream (one processor), its frames
    recorded and replayed from device memory as one hipGraph per replay; the
    VI scanout of config 2 is timed on the rendered frame (scanout_sync).  A
    frame counts as bit-exact when the state digest after it (RDRAM, hidden
    RDRAM and TMEM, computed on the device) equals the oracle's, frame by
    frame, and the final state and, for config 2, the last scanout image
    match byte for byte.  `pending` maps each config to the future of its
    oracle run."""
    rows = {}
    for name, stream in (("c0", rs.config0_stream()), ("c1", rs.config1_stream()), ("c2", rs.config2_stream())):
        frames = len(stream.frames)
        # Per-frame digests: the frames submitted one at a time.
        p = setup_processors(rdp_mi, [stream])[0]
        digests = []
        last_image = None
        for f in stream.frames:
            p.enqueue_commands(f.words)
            if f.vi:
                for r, v in f.vi.items():
                    p.set_vi_register(r, v)
                last_image = p.scanout_sync()
            p.idle()
            digests.append(p.state_digest())
        p.close()
        # Timed: the whole stream recorded once, replayed as a graph.
        p = setup_processors(rdp_mi, [stream])[0]
        p.enqueue_commands(stream.all_words())
        rec = rdp_mi.Recording([p], batched=True)
        if rec.passes < frames:
            raise RuntimeError(f"{name}: the recording holds {rec.passes} passes for {frames} frames")
        rec.replay(1, use_graph=True)
        reps = max(2, min(args.steps, 10))
        ms = float(rec.replay(reps, use_graph=True).mean())
        vi_ms = 0.0
        if name == "c2":
            for r, v in stream.frames[-1].vi.items():
                p.set_vi_register(r, v)
            p.scanout_sync()
            t0 = time.perf_counter()
            for _ in range(frames):
                img = p.scanout_sync()
            vi_ms = (time.perf_counter() - t0) * 1e3 / frames
        state = (p.read_rdram(), p.read_hidden_rdram(), p.read_tmem())
        rec.close()
        p.close()
        ref, ref_digests, ref_images, cpu_s, _ = pending[name].result()
        final_ok = all(np.array_equal(a.view(np.uint8), b.view(np.uint8)) for a, b in zip(state, ref))
        image_ok = name != "c2" or (ref_images and last_image is not None and
                                    np.array_equal(last_image, ref_images[-1]) and np.array_equal(img, ref_images[-1]))
        exact = sum(int(a == b) for a, b in zip(digests, ref_digests)) if final_ok and image_ok else 0
        frame_ms = ms / frames + vi_ms
        rows[name] = {"frames": frames, "triangles": stream.num_triangles,
                      "framebuffer": f"{stream.width}x{stream.height}",
                      "gpu_frames_per_s": round((exact / frames) / (frame_ms / 1e3), 1),
                      "ms_per_frame": round(frame_ms, 4),
                      "render_ms_per_frame": round(ms / frames, 4), "vi_ms_per_frame": round(vi_ms, 4) or None,
                      "bit_exact_frames": f"{exact}/{frames}",
                      "cpu_1core_frames_per_s": round(frames / cpu_s, 3)}
        log(f"[config] {name}: {rows[name]}")
    return rows


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--streams", type=int, default=64)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-plain", action="store_true", help="skip the plain-path comparison (profiling runs)")
    ap.add_argument("--no-serve", action="store_true", help="skip the serving measurement")
    ap.add_argument("--no-configs", action="store_true", help="skip the per-config rows (configs 0-2 on their own)")
    ap.add_argument("--no-upscaled", action="store_true", help="skip the upscaled and super-sampled rows")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the default launch shape (profiling runs: the tuner's trial launches stay out)")
    ap.add_argument("--instances", type=int, default=256)
    ap.add_argument("--peak-frames-per-instance", type=int, default=8)
    ap.add_argument("--serve-seconds", type=float, default=1.0)
    ap.add_argument("--serve-max-frames", type=int, default=4096)
    ap.add_argument("--max-batch", type=int, default=64, help="frames per serving batch at most")
    ap.add_argument("--serve-streams", type=int, default=3, help="submission streams of the server")
    ap.add_argument("--serve-decode-threads", type=int, default=0,
                    help="host decode threads of the server (0: a quarter of the host threads, at most 8)")
    ap.add_argument("--serve-dump", default="", help="write per-request serving times to PREFIX_<load>.npz")
    ap.add_argument("--hbm-budget-mib", type=int, default=2048,
                    help="device bytes one serving batch may use for packed inputs and scratch")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of this many MiB (HBM counter calibration)")
    args = ap.parse_args()

    if args.copy_calibration:
        import torch
        n = args.copy_calibration << 20
        src = torch.empty(n, dtype=torch.uint8, device="cuda").fill_(1)
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes": n, "launches": 3}))
        return

    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (got --gpus {args.gpus}); every stream is independent, "
              "so more GPUs would each run their own batch.", file=sys.stderr)
        sys.exit(2)

    import rdp_mi

    streams = rs.batch_streams(args.streams)
    total_tris = sum(s.num_triangles for s in streams)
    log(f"[bench] {len(streams)} streams, {total_tris} triangles per step")

    threads = cpu_threads()
    refs, cpu_wall, sample_n, sample_s, step_fragments = oracle_results(streams, threads)
    log(f"[bench] oracle: {cpu_wall:.2f}s for {len(streams)} frames on {threads} threads; "
        f"{sample_s:.2f}s for {sample_n} frames on one")
    # The per-config oracle runs (long sequential streams) go on in the
    # background while the GPU is measured.
    from concurrent.futures import ThreadPoolExecutor
    bg = ThreadPoolExecutor(max_workers=3)
    pending = {}
    if not args.no_configs:
        pending = {"c0": bg.submit(oracle_stream, rs.config0_stream(), True),
                   "c1": bg.submit(oracle_stream, rs.config1_stream(), True),
                   "c2": bg.submit(oracle_stream, rs.config2_stream(), True, True)}

    # ---- batched path: one launch per stage per pass round, hipGraph replay
    procs = setup_processors(rdp_mi, streams)
    enqueue_all(procs, streams)
    rec = rdp_mi.Recording(procs, batched=True)
    log(f"[bench] batched recording: {rec.rounds} rounds, {rec.passes} passes")
    if rec.passes < len(streams):
        raise RuntimeError(f"the batched recording holds {rec.passes} passes for {len(streams)} streams")
    if args.no_autotune:
        tuned = {"xcd_chunk": 64, "min_class": 0, "tile_size": 8, "persistent": False, "ms": None}
    else:
        tuned = rec.autotune(3)
        # The best time of each tile size and of the persistent round launch
        # among the candidates.
        for cand in rec.autotune_results():
            key = "best_ms_persistent" if cand["persistent"] else f"best_ms_tile{cand['tile_size']}"
            tuned[key] = round(min(tuned.get(key, 1e9), cand["ms"]), 4)
    log(f"[bench] launch shape: {tuned}")
    # Immediate submissions of the same workload (plain path, serving) use it too.
    rdp_mi.set_launch_config(tuned["xcd_chunk"], tuned["min_class"], tuned["tile_size"], tuned["persistent"])
    rec.replay(max(args.warmup, 1), use_graph=True)
    t0 = time.perf_counter()
    step_ms = rec.replay(args.steps, use_graph=True)
    wall = time.perf_counter() - t0
    exact, mism = exact_frames(procs, refs)
    log(f"[bench] batched: {step_ms.mean():.3f} ms/step (gpu), wall {wall / args.steps * 1e3:.3f} ms/step, "
        f"bit-exact {exact}/{len(streams)} ({mism} mismatched bytes)")
    stage_ms = rec.profile(3)
    launches = rec.profile_raster(3)
    raster_bytes = rec.raster_bytes
    rounds = rec.rounds
    rec_passes = rec.passes
    raster_launches = rec.raster_launches
    rec.close()
    for p in procs:
        p.close()

    # ---- plain path: each pass launched on its own, no graph
    plain_ms, plain_exact = None, 0
    if not args.no_plain:
        procs = setup_processors(rdp_mi, streams)
        enqueue_all(procs, streams)
        plain = rdp_mi.Recording(procs, batched=False)
        if plain.passes != rec_passes:
            raise RuntimeError(f"the plain recording holds {plain.passes} passes, the batched one {rec_passes}")
        plain.replay(max(args.warmup, 1), use_graph=False)
        plain_steps = max(3, args.steps // 5)
        plain_ms = plain.replay(plain_steps, use_graph=False)
        plain_exact, plain_mism = exact_frames(procs, refs)
        log(f"[bench] plain: {plain_ms.mean():.3f} ms/step, bit-exact {plain_exact}/{len(streams)}")
        plain.close()
        for p in procs:
            p.close()

    serve = None if args.no_serve else serving(rdp_mi, streams, refs, args)
    configs = None if args.no_configs else config_rows(rdp_mi, pending, args)
    upscaled = None if args.no_upscaled else upscaled_rows(rdp_mi, streams, args)
    bg.shutdown()

    frames_per_step = len(streams)
    ms_per_step = float(step_ms.mean())
    value = exact / (ms_per_step / 1e3)
    plain_value = plain_exact / (float(plain_ms.mean()) / 1e3) if plain_ms is not None else 0.0

    # The dominant kernel: the raster launch with the longest duration when
    # each launch runs alone (events around it, stages in series, after the
    # timed loop).  Its algorithmic bytes are the framebuffer colour, depth
    # and hidden bits each tile reads and writes once.
    dom = max(launches, key=lambda l: l["ms"])
    achieved = dom["bytes"] / (dom["ms"] / 1e3) / 1e9 if dom["ms"] > 0 else 0.0

    # HBM bytes and SQ instruction counts per launch of the dominant kernel
    # come from the committed rocprofv3 counter runs of this workload
    # (profiles/), not from this run: counters need their own profiled runs.
    traffic, valu, traffic_src = None, None, None
    traffic_path = os.path.join(ROOT, "profiles", "k_raster_batch64_traffic.json")
    if os.path.exists(traffic_path) and args.streams == 64:
        with open(traffic_path) as fh:
            prof = json.load(fh).get(dom["kernel"], {})
        traffic = prof.get("traffic_bytes_per_launch")
        traffic_src = "committed rocprofv3 --pmc FETCH_SIZE/WRITE_SIZE runs (profiles/k_raster_batch64_traffic.json)"
        # VALU issue rate: the committed SQ_INSTS_VALU count per launch over
        # this run's duration of the launch.
        if prof.get("valu_insts_per_launch") and dom["ms"] > 0:
            ops = prof["valu_insts_per_launch"] * 64 / (dom["ms"] / 1e3)
            valu = {"achieved": round(ops / 1e12, 3), "peak": VALU_PEAK_TOPS, "unit": "T lane-ops/s",
                    "frac": round(ops / 1e12 / VALU_PEAK_TOPS, 4),
                    "insts_per_launch": prof["valu_insts_per_launch"],
                    "salu_insts_per_launch": prof.get("salu_insts_per_launch"),
                    "source": "committed SQ_INSTS_VALU count per launch (profiles/k_raster_batch64_traffic.json)"}

    # Algorithmic VALU work of a step: O = sum over render modes of the step's
    # fragments (the oracle's count) x K_mode, the lane-ops per covered
    # fragment of the mode measured on fully covered tiles (committed
    # rocprofv3 calibration, profiles/valu_per_fragment.json).  O over the
    # step time against the VALU peak is the share of the peak that covered
    # fragments account for; valu_issue above counts every issued lane.
    valu_alg = None
    k_path = os.path.join(ROOT, "profiles", "valu_per_fragment.json")
    if os.path.exists(k_path):
        with open(k_path) as fh:
            ktab = json.load(fh)
        ops = sum(step_fragments.get(m, 0) * v["k_lane_ops_per_fragment"] for m, v in ktab.items())
        unpriced = {m: n for m, n in step_fragments.items() if n and m not in ktab}
        rate = ops / (ms_per_step / 1e3)
        valu_alg = {"ops_per_step": round(ops), "achieved": round(rate / 1e12, 3), "peak": VALU_PEAK_TOPS,
                    "unit": "T lane-ops/s", "frac": round(rate / 1e12 / VALU_PEAK_TOPS, 4),
                    "fragments_per_step": step_fragments, "unpriced_fragments": unpriced or None,
                    "k_lane_ops_per_fragment": {m: v["k_lane_ops_per_fragment"] for m, v in ktab.items()},
                    "source": "fragments: the oracle's per-mode count over the step's streams; K: committed "
                              "SQ_INSTS_VALU calibration on fully covered tiles (profiles/valu_per_fragment.json)"}

    cpu = None
    if not args.no_cpu_baseline:
        cpu = {"value": round(frames_per_step / cpu_wall, 3), "unit": "frames/s", "cores": threads, "kind": "port",
               "sample": f"all {frames_per_step} batch streams, one frame each, through the C oracle, one stream "
                         f"per thread on {threads} host threads",
               "one_core_value": round(sample_n / sample_s, 3),
               "one_core_sample": f"the first {sample_n} batch streams in sequence on one thread",
               "host_cpu": cpu_model(), "host_threads_available": os.cpu_count()}

    out = {
        "metric": "bit-exact RDP frames/s, 64-stream batch vs plain path; p99 frame latency",
        "value": round(value, 3),
        "unit": "frames/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32",
        "data": "synthetic",
        "config": {"workload": f"batched form: {frames_per_step} seeded streams (8 MB RDRAM each), "
                               f"{total_tris} triangles per step, 320x240/640x480 at 16/32 bpp, "
                               "modes drawn from the shaded, textured and 2-cycle CI/LOD/fog/AA workloads; "
                               "shaded streams of more than 512 triangles use 2-200 px triangles rather than "
                               "full-screen ones",
                   "streams": frames_per_step, "triangles_per_step": total_tris},
        "bit_exact_frames": f"{exact}/{frames_per_step}",
        "mismatched_bytes": mism,
        "p50_frame_latency_ms": None if serve is None else serve["loads"][1]["p50_ms"],
        "p99_frame_latency_ms": None if serve is None else serve["loads"][1]["p99_ms"],
        "latency_note": "serving form, Poisson arrivals at 50% of peak offered load, completion minus arrival",
        "plain_path": None if plain_ms is None else {
            "value": round(plain_value, 3), "unit": "frames/s", "ms_per_step": round(float(plain_ms.mean()), 4),
            "bit_exact_frames": f"{plain_exact}/{frames_per_step}"},
        "speedup_vs_plain": round(value / plain_value, 3) if plain_value > 0 else None,
        "autotuned": tuned,
        "stage_ms_per_step": {"setup": round(float(stage_ms[0]), 4), "tmem": round(float(stage_ms[1]), 4),
                              "bin": round(float(stage_ms[2]), 4), "raster": round(float(stage_ms[3]), 4)},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": traffic, "traffic_source": traffic_src,
                     "kernel": dom["kernel"], "bytes_per_launch": dom["bytes"], "tiles_per_launch": dom["tiles"],
                     "ms_per_launch": round(dom["ms"], 4), "timing": "the launch alone, HIP events on its stream",
                     "valu_issue": valu,
                     "valu_algorithmic": valu_alg,
                     "launches": [{k: (round(v, 4) if isinstance(v, float) else v) for k, v in l.items()}
                                  for l in launches]},
        "cpu_baseline": cpu,
        "configs": configs,
        "serving": serve,
        "upscaled": upscaled,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
