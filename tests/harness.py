"""Shared helpers for the parity tests: run one stream through the oracle and
through the MI355X library, then compare RDRAM, hidden RDRAM and TMEM."""
from __future__ import annotations

import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for sub in ("oracle", "parallel-rdp_amd"):
    p = os.path.join(ROOT, sub)
    if p not in sys.path:
        sys.path.insert(0, p)

import rdp_streams as rs  # noqa: E402


def run_oracle(stream, frames=None, scanouts=None):
    """Renders the stream on the oracle.  With a list for `scanouts`, frames
    that carry VI registers are scanned out after their commands and the
    images appended."""
    import oracle
    o = oracle.Oracle(rs.RDRAM_SIZE)
    if stream.rdram_init is not None:
        o.rdram[:] = stream.rdram_init
    o.hidden[:] = stream.hidden_init
    for f in (stream.frames if frames is None else stream.frames[:frames]):
        o.process(f.words)
        if scanouts is not None and f.vi:
            for r, v in f.vi.items():
                o.set_vi_register(r, v)
            scanouts.append(o.scanout())
    o.flush()
    out = (o.rdram.copy(), o.hidden.copy(), o.tmem.copy(), o.fragments)
    o.close()
    return out


def run_gpu(stream, frames=None, proc=None, scanouts=None):
    import rdp_mi
    p = proc or rdp_mi.CommandProcessor(rs.RDRAM_SIZE)
    p.write_rdram(stream.rdram_init if stream.rdram_init is not None else np.zeros(rs.RDRAM_SIZE // 4, np.uint32))
    p.write_hidden_rdram(np.full(rs.RDRAM_SIZE // 2, stream.hidden_init, np.uint8))
    for f in (stream.frames if frames is None else stream.frames[:frames]):
        p.enqueue_commands(f.words)
        if scanouts is not None and f.vi:
            for r, v in f.vi.items():
                p.set_vi_register(r, v)
            scanouts.append(p.scanout_sync())
    p.idle()
    out = (p.read_rdram(), p.read_hidden_rdram(), p.read_tmem())
    if proc is None:
        p.close()
    return out


def mismatch_report(a, b, name):
    a8 = a.view(np.uint8)
    b8 = b.view(np.uint8)
    diff = np.nonzero(a8 != b8)[0]
    if diff.size == 0:
        return 0, f"{name}: identical"
    return int(diff.size), f"{name}: {diff.size} mismatched bytes, first at {hex(int(diff[0]))}"


def compare(gpu, ref):
    """Returns (total mismatched bytes, report lines)."""
    total = 0
    lines = []
    for name, g, r in (("rdram", gpu[0], ref[0]), ("hidden", gpu[1], ref[1]), ("tmem", gpu[2], ref[2])):
        n, line = mismatch_report(g, r, name)
        total += n
        lines.append(line)
    return total, lines


def _oracle_state(stream):
    rd, hid, tm, _ = run_oracle(stream)
    return rd, hid, tm


def run_oracle_many(streams, workers=8):
    """run_oracle over many streams in worker processes (spawned: they load
    only the oracle, never the device); results in stream order, without the
    fragment counts."""
    import concurrent.futures
    import multiprocessing
    ctx = multiprocessing.get_context("spawn")
    with concurrent.futures.ProcessPoolExecutor(max_workers=workers, mp_context=ctx) as pool:
        return list(pool.map(_oracle_state, streams))


def oracle_frame_digests(stream):
    """The oracle's state digest after each frame of the stream."""
    import oracle
    o = oracle.Oracle(rs.RDRAM_SIZE)
    if stream.rdram_init is not None:
        o.rdram[:] = stream.rdram_init
    o.hidden[:] = stream.hidden_init
    out = []
    for f in stream.frames:
        o.process(f.words)
        o.flush()
        out.append(o.digest())
    o.close()
    return out
