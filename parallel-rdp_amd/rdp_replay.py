#!/usr/bin/env python3
"""Replay an RDPDUMP2 capture on the MI355X and write what it shows as PNG
images: the VI scanout of every frame and, at chosen points, the colour,
depth and coverage views of the framebuffer the capture last selected (the
capture inspection of the reference's rdp_replayer.cpp:246-422 and
replay_until, :637; its interactive window is not reproduced).

    python parallel-rdp_amd/rdp_replay.py CAPTURE --out DIR [--every-draws N]
        [--views color,depth,coverage] [--frames A:B] [--no-scanout]

Per frame F (counted from 0 at each EndFrame):
    DIR/frame_F_scanout.png             the scanout (blend_previous_frame on,
                                        upscale_deinterlacing off, as the
                                        reference's replay driver scans out)
    DIR/frame_F_<view>.png              the views at the end of the frame
    DIR/frame_F_draw_D_<view>.png       with --every-draws N: the views after
                                        every N-th draw call of the frame
The colour image follows SetColorImage and the depth image SetMaskImage
commands as the capture replays (rdp_replayer.cpp:395-421).
"""
from __future__ import annotations

import argparse
import os
import struct
import sys
import zlib

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
if HERE not in sys.path:
    sys.path.insert(0, HERE)

import rdp_mi  # noqa: E402

VIEWS = {"color": rdp_mi.CommandProcessor.VIEW_COLOR, "depth": rdp_mi.CommandProcessor.VIEW_DEPTH,
         "coverage": rdp_mi.CommandProcessor.VIEW_COVERAGE}


def png_bytes(rgba: np.ndarray) -> bytes:
    """PNG (8-bit RGBA, no filtering) of a (height, width) array of RGBA8
    words with R in the low byte."""
    h, w = rgba.shape
    raw = np.ascontiguousarray(rgba, dtype="<u4").view(np.uint8).reshape(h, w * 4)
    rows = np.concatenate([np.zeros((h, 1), np.uint8), raw], axis=1).tobytes()

    def chunk(tag: bytes, data: bytes) -> bytes:
        return struct.pack(">I", len(data)) + tag + data + struct.pack(">I", zlib.crc32(tag + data) & 0xFFFFFFFF)

    return (b"\x89PNG\r\n\x1a\n" + chunk(b"IHDR", struct.pack(">IIBBBBB", w, h, 8, 6, 0, 0, 0)) +
            chunk(b"IDAT", zlib.compress(rows, 6)) + chunk(b"IEND", b""))


def write_png(path: str, rgba: np.ndarray):
    with open(path, "wb") as fh:
        fh.write(png_bytes(rgba))


class FramebufferTracker:
    """The colour and depth images the capture last selected
    (rdp_replayer.cpp:395-421)."""

    def __init__(self):
        self.fb_addr = self.fb_size = self.fb_width = self.depth_addr = 0

    def command(self, op: int, words: np.ndarray):
        if op == 0x3F:  # SetColorImage
            self.fb_size = (int(words[0]) >> 19) & 3
            self.fb_width = (int(words[0]) & 0x3FF) + 1
            self.fb_addr = int(words[1]) & 0xFFFFFF
        elif op == 0x3E:  # SetMaskImage
            self.depth_addr = int(words[1]) & 0xFFFFFF


def replay(capture: str, out: str, every_draws: int = 0, views=("color", "depth", "coverage"),
           frames=None, scanout: bool = True) -> dict:
    os.makedirs(out, exist_ok=True)
    fb = FramebufferTracker()
    frame = draw = written = 0
    opts = rdp_mi.ScanoutOptions.default(blend_previous_frame=1, upscale_deinterlacing=0)

    def wanted():
        return frames is None or frames[0] <= frame < frames[1]

    def dump_views(tag):
        nonlocal written
        if not fb.fb_width:
            return
        for name in views:
            img = proc.framebuffer_view(VIEWS[name], fb.fb_addr, fb.fb_size, fb.depth_addr, fb.fb_width)
            write_png(os.path.join(out, f"frame_{frame}_{tag}{name}.png"), img)
            written += 1

    with rdp_mi.DumpPlayer(capture) as pl:
        proc = rdp_mi.CommandProcessor(pl.rdram_size)
        proc.write_hidden_rdram(np.zeros(pl.rdram_size // 2, np.uint8))  # the player's shadows start zeroed
        try:
            while True:
                ev = pl.next()
                if ev is None:
                    break
                kind = ev.kind
                if kind == rdp_mi.DUMP_RDP_COMMAND:
                    fb.command(ev.op, pl.words())
                pl.apply(proc)
                if kind == rdp_mi.DUMP_RDP_COMMAND and ev.op in rdp_mi.DRAW_OPS:
                    draw += 1
                    if every_draws and draw % every_draws == 0 and wanted():
                        dump_views(f"draw_{draw}_")
                elif kind == rdp_mi.DUMP_END_FRAME:
                    if wanted():
                        if scanout:
                            img = proc.scanout_sync(opts)
                            if img.size:
                                write_png(os.path.join(out, f"frame_{frame}_scanout.png"), img)
                                written += 1
                        dump_views("")
                    frame += 1
                    draw = 0
            proc.idle()
        finally:
            proc.close()
    return {"frames": frame, "images": written}


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("capture")
    ap.add_argument("--out", required=True)
    ap.add_argument("--every-draws", type=int, default=0)
    ap.add_argument("--views", default="color,depth,coverage")
    ap.add_argument("--frames", default="", help="A:B, frames A .. B-1 only")
    ap.add_argument("--no-scanout", action="store_true")
    a = ap.parse_args(argv)
    views = [v for v in a.views.split(",") if v]
    for v in views:
        if v not in VIEWS:
            ap.error(f"unknown view {v}")
    frames = tuple(int(x) for x in a.frames.split(":")) if a.frames else None
    res = replay(a.capture, a.out, a.every_draws, views, frames, not a.no_scanout)
    print(res)


if __name__ == "__main__":
    main()
