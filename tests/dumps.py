"""RDPDUMP2 captures for the dump tests: writing a capture of a command
stream the way the reference's capture hook does, an independent reader of
the format, and a replay that drives the oracle.

Capture (CommandProcessor::enqueue_command / set_vi_register / scanout,
rdp_device.cpp:874-897, 970-975, 1087-1092): the first command of a command
list records the RDRAM and hidden RDRAM blocks changed since the last record
(the state after all earlier work), every command is recorded except SyncFull,
which becomes SignalComplete and ends the list, VI registers are recorded,
and a scanout records the changed blocks and EndFrame.  The oracle supplies
the RDRAM state at those points.
"""
from __future__ import annotations

import gzip
import os
import shutil
import struct
import tempfile

import numpy as np

import rdp_streams as rs

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden", "dumps")

# Record kinds (rdp_dump.cpp:28-40).
UPDATE_DRAM, RDP_COMMAND, SET_VI_REGISTER, END_FRAME, SIGNAL_COMPLETE, EOF = 1, 2, 3, 4, 5, 6
UPDATE_DRAM_FLUSH, UPDATE_HIDDEN_DRAM, UPDATE_HIDDEN_DRAM_FLUSH = 7, 8, 9
SYNC_FULL = 0x29
DRAW_OPS = frozenset((0x08, 0x09, 0x0a, 0x0b, 0x0c, 0x0d, 0x0e, 0x0f, 0x24, 0x25, 0x36))


def commands(words):
    """Split a word stream into complete commands (undefined opcodes skipped)."""
    import rdp_mi
    i = 0
    out = []
    while i < words.size:
        n = int(rdp_mi.COMMAND_LENGTHS[(int(words[i]) >> 24) & 63])
        if n == 0:
            i += 1
            continue
        if i + n > words.size:
            break
        out.append(words[i:i + n])
        i += n
    return out


def write_capture(stream, path):
    """Capture of `stream` (every frame; frames with VI registers are scanned
    out) through the library's RDPDUMP2 writer, RDRAM states from the oracle."""
    import oracle
    import rdp_mi
    o = oracle.Oracle(rs.RDRAM_SIZE)
    if stream.rdram_init is not None:
        o.rdram[:] = stream.rdram_init
    o.hidden[:] = stream.hidden_init
    try:
        with rdp_mi.DumpWriter(path, rs.RDRAM_SIZE, rs.RDRAM_SIZE // 2) as w:
            in_list = False
            for f in stream.frames:
                for cmd in commands(f.words):
                    if not in_list:
                        o.flush()
                        w.flush_rdram(o.rdram)
                        w.flush_hidden_rdram(o.hidden)
                        in_list = True
                    o.process(cmd)
                    if (int(cmd[0]) >> 24) & 63 == SYNC_FULL:
                        w.signal_complete()
                        in_list = False
                    else:
                        w.command(cmd)
                if f.vi:
                    for reg, val in f.vi.items():
                        w.set_vi_register(reg, val)
                    o.flush()
                    w.flush_rdram(o.rdram)
                    w.flush_hidden_rdram(o.hidden)
                    w.end_frame()
    finally:
        o.close()


def read_records(data: bytes):
    """Independent reader of RDPDUMP2 (the format of rdp_dump.cpp:43-200):
    returns (rdram_size, hidden_size, records); each record is a tuple
    (kind, payload)."""
    if data[:8] != b"RDPDUMP2":
        raise ValueError("not an RDPDUMP2 capture")
    rdram_size, hidden_size = struct.unpack_from("<II", data, 8)
    pos = 16
    recs = []

    def u32():
        nonlocal pos
        if pos + 4 > len(data):
            raise ValueError("truncated record")
        v = struct.unpack_from("<I", data, pos)[0]
        pos += 4
        return v

    while pos < len(data):
        kind = u32()
        if kind == EOF:
            recs.append((EOF, None))
            break
        if kind in (END_FRAME, SIGNAL_COMPLETE, UPDATE_DRAM_FLUSH, UPDATE_HIDDEN_DRAM_FLUSH):
            recs.append((kind, None))
        elif kind == SET_VI_REGISTER:
            recs.append((kind, (u32(), u32())))
        elif kind == RDP_COMMAND:
            op, n = u32(), u32()
            if pos + 4 * n > len(data):
                raise ValueError("truncated command")
            words = np.frombuffer(data, dtype="<u4", count=n, offset=pos).astype(np.uint32)
            pos += 4 * n
            recs.append((kind, (op, words)))
        elif kind in (UPDATE_DRAM, UPDATE_HIDDEN_DRAM):
            off, size = u32(), u32()
            if pos + size > len(data):
                raise ValueError("truncated block")
            recs.append((kind, (off, data[pos:pos + size])))
            pos += size
        else:
            raise ValueError(f"unknown record kind {kind}")
    return rdram_size, hidden_size, recs


def golden_path(name):
    """A committed capture, decompressed to a temporary file (the files are
    kept gzip-compressed in the repository)."""
    src = os.path.join(GOLDEN, name + ".rdp.gz")
    fd, dst = tempfile.mkstemp(suffix=".rdp", prefix=name + "_")
    with os.fdopen(fd, "wb") as out, gzip.open(src, "rb") as inp:
        shutil.copyfileobj(inp, out)
    return dst


GOLDEN_NAMES = ("c0_shaded_64", "c2_two_cycle_320x240_vi", "b3_two_cycle_rgba8888", "render_to_texture")


def golden_streams():
    """The streams the committed captures were written from."""
    import test_gpu_rounds as rounds
    return {
        "c0_shaded_64": rs.config0_stream(num_triangles=64),
        "c2_two_cycle_320x240_vi": rs.config2_stream(frames=2, prims_per_frame=200, width=320, height=240,
                                                     num_textures=4),
        "b3_two_cycle_rgba8888": rs.batch_stream(3, max_tris=200),
        "render_to_texture": _small_rdram(rounds.texture_stream(400, True), [(0x1FF000, 0x202000), (0x300000, 0x302000)]),
    }


def _small_rdram(stream, ranges):
    """The stream with its random RDRAM kept only in the byte ranges [lo, hi)
    and hidden RDRAM zeroed, so its capture stays small."""
    rdram = np.zeros(rs.RDRAM_SIZE // 4, np.uint32)
    for lo, hi in ranges:
        rdram[lo // 4:hi // 4] = stream.rdram_init[lo // 4:hi // 4]
    stream.rdram_init = rdram
    stream.hidden_init = 0
    return stream


class OracleReplay:
    """Drives the oracle with capture records as the replay driver drives a
    processor (replayer_driver_parallel.cpp:124-205): commands are processed,
    SignalComplete ends the pass, block flushes replace RDRAM / hidden RDRAM."""

    def __init__(self, rdram_size):
        import oracle
        self.o = oracle.Oracle(rdram_size)
        self.o.hidden[:] = 0  # the player's shadows start zeroed

    def apply(self, kind, payload, shadow_rdram, shadow_hidden):
        o = self.o
        if kind == RDP_COMMAND:
            o.process(payload[1])
        elif kind == SET_VI_REGISTER:
            o.set_vi_register(*payload)
        elif kind == SIGNAL_COMPLETE:
            o.flush()
        elif kind == UPDATE_DRAM_FLUSH:
            o.flush()
            o.rdram[:] = shadow_rdram
        elif kind == UPDATE_HIDDEN_DRAM_FLUSH:
            o.flush()
            o.hidden[:] = shadow_hidden

    def state(self):
        self.o.flush()
        return self.o.rdram.copy(), self.o.hidden.copy(), self.o.tmem.copy()

    def close(self):
        self.o.close()
