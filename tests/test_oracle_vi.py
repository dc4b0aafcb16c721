"""VI scanout oracle: the reference's per-scanline-xh known answer
(vi_conformance.cpp:163-267), closed forms of the divot and fetch filters,
and frame-control behaviour (blank / invalid frames, output size)."""
import numpy as np
import pytest

import rdp_streams as rs

H_OFF, V_OFF = 108, 34            # VI_H_OFFSET_NTSC, VI_V_OFFSET_NTSC
RGBA8888, RGBA5551 = 3, 2
AA_RESAMP_ONLY, AA_REPLICATE = 2 << 8, 3 << 8


def start_reg(a, b):
    return ((a & 0x3FF) << 16) | (b & 0x3FF)


def scale_reg(scale, bias):
    return ((bias & 0xFFF) << 16) | (scale & 0xFFF)


def default_vi(o, fmt=RGBA5551, aa=AA_REPLICATE, extra=0, pal=False):
    """set_default_vi_registers (vi_conformance.cpp:62-91), including its
    end_frame with an empty HStart that clears persistent state."""
    o.set_vi_register(rs.VI_CONTROL, aa | fmt | extra)
    o.set_vi_register(rs.VI_ORIGIN, 567123)
    o.set_vi_register(rs.VI_WIDTH, 100)
    o.set_vi_register(rs.VI_V_SYNC, 625 if pal else 525)
    voff = 44 if pal else V_OFF
    o.set_vi_register(rs.VI_V_START, start_reg(voff, voff + 224 * 2))
    o.set_vi_register(rs.VI_X_SCALE, scale_reg(1024, 0))
    o.set_vi_register(rs.VI_Y_SCALE, scale_reg(1024, 0))
    o.set_vi_register(rs.VI_H_START, start_reg(640, 0))
    o.scanout(suite_options())
    hoff = 128 if pal else H_OFF
    o.set_vi_register(rs.VI_H_START, start_reg(hoff, hoff + 640))


def suite_options(**kw):
    import oracle
    return oracle.ScanoutOptions.default(blend_previous_frame=1, upscale_deinterlacing=0, **kw)


def per_scanline_reference(crop=None, sf=1):
    """The expected image of run_per_scanline_xh_vi (red channel only); sf
    is its scale_factor (2 for the upscaled variant): the image is sf times
    larger and a sample at x_lo of the larger image reads native pixel
    x_lo / sf."""
    ref = np.zeros((240 * sf, 640 * sf), dtype=np.int64)

    def region(y0, y1, h_start, h_end, left_clamp, right_clamp, x_add, x_start):
        x_base = h_start
        if not left_clamp:
            h_start += 8
        if not right_clamp:
            h_end -= 7
        x_begin, x_end = max(h_start, 0) * sf, min(h_end, 640) * sf
        for x in range(x_begin, x_end):
            sx = (x - x_base * sf) * x_add + x_start * sf
            frac = (sx >> 5) & 31
            lo = sx >> 10
            hi = lo + 1
            lo, hi = lo // sf, hi // sf
            ref[y0 * sf:y1 * sf, x] = (lo * (32 - frac) + hi * frac + 16) >> 5

    region(20, 50, 0, 320, False, False, 256, 0)
    region(50, 100, 0, 640, False, False, 240, 200)
    region(100, 150, -8, 648, True, True, 220, 400)
    region(150, 200, 8, 640, False, True, 210, 600)
    if crop:
        l, r, t, b = (c * sf for c in crop)
        ref = ref[t:240 * sf - b, l:640 * sf - r]
    return ref


@pytest.mark.parametrize("crop,factor", [(None, 1), ((9, 11, 10, 12), 1), (None, 2), ((9, 11, 10, 12), 2)])
def test_per_scanline_xh_known_answer(oracle_mod, crop, factor):
    """factor 2: the reference's upscaled variant (state.gpu_scaled); the
    framebuffer is written to native RDRAM only and reaches the sample
    planes through the scanout's domain update."""
    o = oracle_mod.Oracle(rs.RDRAM_SIZE)
    if factor > 1:
        o.enable_upscaling(factor)
    default_vi(o, fmt=RGBA8888, aa=AA_RESAMP_ONLY)
    fb = np.zeros(240 * 200, dtype=np.uint32)
    fb[:] = (np.tile(np.arange(200, dtype=np.uint32), 240) << 24)
    o.rdram[4096 // 4: 4096 // 4 + fb.size] = fb
    o.set_vi_register(rs.VI_ORIGIN, 4096)
    o.set_vi_register(rs.VI_WIDTH, 200)
    o.set_vi_register(rs.VI_V_START, start_reg(V_OFF + 20 * 2, V_OFF + 200 * 2))
    o.begin_vi_register_per_scanline(3)
    for line, hs, xs in ((V_OFF, start_reg(H_OFF, H_OFF + 320), scale_reg(256, 0)),
                         (V_OFF + 100, start_reg(H_OFF, H_OFF + 640), scale_reg(240, 200)),
                         (V_OFF + 200, start_reg(H_OFF - 8, H_OFF + 648), scale_reg(220, 400)),
                         (V_OFF + 300, start_reg(H_OFF + 8, H_OFF + 648), scale_reg(210, 600))):
        o.set_vi_register_for_scanline(1, hs)
        o.set_vi_register_for_scanline(2, xs)
        o.latch_vi_register_for_scanline(line)
    o.end_vi_register_per_scanline()
    kw = {}
    if crop:
        kw = dict(crop_enable=1, crop_left=crop[0], crop_right=crop[1], crop_top=crop[2], crop_bottom=crop[3])
    img = o.scanout(suite_options(**kw))
    ref = per_scanline_reference(crop, factor)
    assert img.shape == ref.shape
    rgb = np.stack([img & 0xFF, (img >> 8) & 0xFF, (img >> 16) & 0xFF], axis=-1).astype(np.int64)
    assert np.array_equal(rgb[:, :, 0], ref)
    assert not rgb[:, :, 1:].any()
    o.close()


def test_blank_and_invalid_frames(oracle_mod):
    o = oracle_mod.Oracle(rs.RDRAM_SIZE)
    # Origin 0: nothing.
    assert o.scanout().shape == (0, 0)
    default_vi(o)
    img = o.scanout()
    assert img.shape == (240, 640)        # fields come out progressive by default
    assert o.scanout(suite_options()).shape == (240, 640)
    # Blank type twice in a row: the second gives 0x0.
    o.set_vi_register(rs.VI_CONTROL, 0)
    assert o.scanout().shape == (240, 640)
    assert o.scanout().shape == (0, 0)
    # PAL with serrate and weave output: 576 lines.
    default_vi(o, pal=True, extra=1 << 6)
    assert o.scanout(suite_options()).shape == (576, 640)
    o.close()


def second_extremes(mid, neigh):
    """Second smallest and second largest of the neighbours with full
    coverage together with the centre pixel counted twice, per channel (the
    running min/max of vi_fetch.frag:42-53 starts all four trackers at the
    centre value)."""
    srt = sorted([mid, mid] + [c for c, a in neigh if a == 7])
    return srt[1], srt[-2]


def test_aa_and_divot_filters_closed_form(oracle_mod):
    """RESAMP_EXTRA at 1:1 scale (bilinear with zero fractions) with divot:
    the partially covered pixel is pulled toward the second-lowest/highest of
    its six neighbours by (7 - coverage)/8 (vi_fetch.frag:73-125), and the
    three pixels around it become per-channel three-tap medians
    (vi_divot.frag:54-92)."""
    rng = np.random.default_rng(5)
    o = oracle_mod.Oracle(rs.RDRAM_SIZE)
    default_vi(o, fmt=RGBA8888, aa=1 << 8, extra=1 << 4)
    o.set_vi_register(rs.VI_ORIGIN, 8192)
    o.set_vi_register(rs.VI_WIDTH, 640)
    px = rng.integers(0, 256, size=(240, 640, 3))
    cov = np.full((240, 640), 7)
    cov[100, 300] = 3
    cov[99, 299] = 2   # a neighbour without full coverage is ignored
    words = (px[..., 0] << 24) | (px[..., 1] << 16) | (px[..., 2] << 8) | (cov << 5)
    o.rdram[8192 // 4: 8192 // 4 + words.size] = words.ravel().astype(np.uint32)
    img = o.scanout()
    got = np.stack([img & 0xFF, (img >> 8) & 0xFF, (img >> 16) & 0xFF], axis=-1).astype(np.int64)

    def aa_value(y, x):
        out = []
        for ch in range(3):
            mid = int(px[y, x, ch])
            neigh = [(int(px[y + dy, x + dx, ch]), int(cov[y + dy, x + dx]))
                     for dx, dy in ((-1, -1), (1, -1), (-2, 0), (2, 0), (-1, 1), (1, 1))]
            lo2, hi2 = second_extremes(mid, neigh)
            off = (lo2 + hi2 - 2 * mid) & 0xFFFFFFFF
            out.append((mid + (((off * (7 - int(cov[y, x])) + 4) & 0xFFFFFFFF) >> 3)) & 0xFF)
        return np.array(out)

    def a_value(y, x):
        return aa_value(y, x) if cov[y, x] != 7 else px[y, x]

    for x in (299, 300, 301):
        trio = np.stack([a_value(100, x - 1), a_value(100, x), a_value(100, x + 1)])
        exp = np.median(trio, axis=0).astype(np.int64)
        assert np.array_equal(got[100, x], exp), (x, got[100, x], exp)
    assert np.array_equal(got[50, 100:200], px[50, 100:200])
    o.close()
