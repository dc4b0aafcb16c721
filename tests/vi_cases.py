"""Random VI register / option sequences for scanout parity tests."""
from __future__ import annotations

import numpy as np

import rdp_streams as rs


def start_reg(a, b):
    return ((int(a) & 0x3FF) << 16) | (int(b) & 0x3FF)


def scale_reg(scale, bias):
    return ((int(bias) & 0xFFF) << 16) | (int(scale) & 0xFFF)


OPTION_FIELDS = ("crop_overscan_pixels", "crop_left", "crop_right", "crop_top", "crop_bottom", "crop_enable",
                 "persist_frame_on_invalid_input", "blend_previous_frame", "upscale_deinterlacing", "vi_aa",
                 "vi_scale", "vi_serrate", "vi_dither_filter", "vi_divot_filter", "vi_gamma_dither")


def random_frame(rng, pal_bias=0.3):
    """One frame's VI registers and scanout options."""
    pal = rng.random() < pal_bias
    hoff, voff = (128, 44) if pal else (108, 34)
    fmt = int(rng.choice([0, 1, 2, 3], p=[0.08, 0.04, 0.44, 0.44]))
    control = fmt | (int(rng.integers(0, 4)) << 8)
    for bit, p in ((1 << 2, 0.3), (1 << 3, 0.4), (1 << 4, 0.5), (1 << 6, 0.3), (1 << 16, 0.5)):
        if rng.random() < p:
            control |= bit
    width = int(rng.choice([160, 320, 640, int(rng.integers(1, 1024))]))
    origin = int(rng.integers(0, rs.RDRAM_SIZE)) & ~3 if rng.random() < 0.95 else 0
    if rng.random() < 0.8:
        h0 = hoff + int(rng.integers(-16, 24))
        h1 = h0 + int(rng.choice([640, 320, int(rng.integers(0, 700))]))
    else:
        h0, h1 = int(rng.integers(0, 1024)), int(rng.integers(0, 1024))
    if rng.random() < 0.8:
        v0 = voff + int(rng.integers(-12, 20))
        v1 = v0 + int(rng.choice([448, 480, 576 if pal else 474, int(rng.integers(0, 600))]))
    else:
        v0, v1 = int(rng.integers(0, 1024)), int(rng.integers(0, 1024))
    xs = int(rng.choice([512, 1024, 640, int(rng.integers(1, 4096))]))
    ys = int(rng.choice([512, 1024, 1000, int(rng.integers(1, 4096))]))
    regs = {
        rs.VI_CONTROL: control,
        rs.VI_ORIGIN: origin,
        rs.VI_WIDTH: width,
        rs.VI_V_SYNC: 625 if pal else 525,
        rs.VI_V_START: start_reg(v0, v1),
        rs.VI_H_START: start_reg(h0, h1),
        rs.VI_X_SCALE: scale_reg(xs, int(rng.integers(0, 4096)) if rng.random() < 0.5 else 0),
        rs.VI_Y_SCALE: scale_reg(ys, int(rng.integers(0, 4096)) if rng.random() < 0.5 else 0),
        rs.VI_V_CURRENT_LINE: int(rng.integers(0, 2)),
    }
    opts = dict(upscale_deinterlacing=int(rng.random() < 0.6), blend_previous_frame=int(rng.random() < 0.4),
                persist_frame_on_invalid_input=int(rng.random() < 0.2),
                vi_aa=int(rng.random() < 0.9), vi_scale=int(rng.random() < 0.9), vi_serrate=int(rng.random() < 0.9),
                vi_dither_filter=int(rng.random() < 0.9), vi_divot_filter=int(rng.random() < 0.9),
                vi_gamma_dither=int(rng.random() < 0.9))
    if rng.random() < 0.2:
        opts.update(crop_enable=1, crop_left=int(rng.integers(0, 40)), crop_right=int(rng.integers(0, 40)),
                    crop_top=int(rng.integers(0, 30)), crop_bottom=int(rng.integers(0, 30)))
    elif rng.random() < 0.2:
        opts.update(crop_overscan_pixels=int(rng.integers(0, 24)))
    per_line = None
    if rng.random() < 0.25:
        per_line = []
        line = voff
        for _ in range(int(rng.integers(1, 5))):
            line += int(rng.integers(1, 120))
            per_line.append((line, start_reg(hoff + int(rng.integers(-16, 16)), hoff + int(rng.integers(200, 660))),
                             scale_reg(int(rng.integers(128, 2048)), int(rng.integers(0, 1024)))))
    return regs, opts, per_line


def random_case(seed, frames=4):
    rng = np.random.default_rng(seed)
    rdram = rng.integers(0, 1 << 32, size=rs.RDRAM_SIZE // 4, dtype=np.uint64).astype(np.uint32)
    hidden = rng.integers(0, 4, size=rs.RDRAM_SIZE // 2).astype(np.uint8)
    return rdram, hidden, [random_frame(rng) for _ in range(frames)]


def apply_frame(target, regs, per_line):
    for r, v in regs.items():
        target.set_vi_register(r, v)
    if per_line:
        target.begin_vi_register_per_scanline(3)
        for line, hs, xs in per_line:
            target.set_vi_register_for_scanline(1, hs)
            target.set_vi_register_for_scanline(2, xs)
            target.latch_vi_register_for_scanline(line)
        target.end_vi_register_per_scanline()
