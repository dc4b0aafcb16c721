"""Upscaled rendering in the oracle: the S-times render pass
(rdp_renderer.cpp:2287-2342 with span_setup.comp, interpolation.h and
shading.h at SCALING_LOG2 > 0), the sample planes
(memory_interfacing.h:55-157, 286-362) and the domain updates
(update_upscaled_domain_pre/post.comp, rdp_renderer.cpp:1936-2033).

Pinned by what the geometry fixes without the oracle's formulas:
pixel-aligned flat rectangles and copy-mode blits cover whole pixels, so
every sample plane must equal the native framebuffer; a smooth gradient's
first sample sits where the native pixel is evaluated, so plane 0 must
agree with native within rounding; and the domain rules fix which memory
reaches the planes (host writes inside a drawn framebuffer, native-only
passes) and which does not (outside it, or memory the upscaled pass owns).
"""
import os
import subprocess

import numpy as np
import pytest

import rdp_streams as rs
import test_oracle_vi as tvi
from rdp_streams import CommandList, OtherModes

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE = os.path.join(ROOT, "oracle")

W, H = 320, 240
COLOR = rs.COLOR_ADDR
PRIM = ((rs.MULADD_ZERO, rs.MULSUB_ZERO, rs.MUL_ZERO, rs.RGB_PRIM), (rs.A_ZERO, rs.A_ZERO, rs.AMUL_ZERO, rs.A_PRIM))
SHADE = ((rs.MULADD_ZERO, rs.MULSUB_ZERO, rs.MUL_ZERO, rs.RGB_SHADE),
         (rs.A_ZERO, rs.A_ZERO, rs.AMUL_ZERO, rs.A_SHADE))
TEX0 = ((rs.MULADD_ZERO, rs.MULSUB_ZERO, rs.MUL_ZERO, rs.RGB_TEXEL0),
        (rs.A_ZERO, rs.A_ZERO, rs.AMUL_ZERO, rs.A_TEXEL0))
ONE_CYCLE = OtherModes(cycle_type=0, blend=((0, 0, 1, 0), (0, 0, 1, 0)))


def _bind(cl, width=W, height=H, size=rs.SIZ_16, fmt=rs.FMT_RGBA, addr=COLOR):
    # Texture rectangles render at scale (the reference's default quirk
    # draws them at native resolution).
    cl.set_quirks(native_resolution_tex_rect=False)
    cl.set_scissor(0, 0, width << 2, height << 2)
    cl.set_color_image(addr, width, size, fmt)
    cl.set_mask_image(rs.DEPTH_ADDR)


def _maker(mod, lib_path=None):
    def make(factor):
        o = mod.Oracle(rs.RDRAM_SIZE, lib_path=lib_path)
        if factor > 1:
            o.enable_upscaling(factor)
        return o
    return make


def _fb(words_u32, size, addr=COLOR, width=W, height=H):
    """The framebuffer as (height, width) pixels in N64 order."""
    if size == rs.SIZ_16:
        return rs.rdram_read16(words_u32, addr, width * height).reshape(height, width)
    base = addr // 4
    return np.asarray(words_u32[base:base + width * height]).reshape(height, width)


def _hidden(hidden_u8, size, addr=COLOR, width=W, height=H):
    n = width * height * (1 if size == rs.SIZ_16 else 2)
    return np.asarray(hidden_u8[addr // 2: addr // 2 + n])


def _planes_equal_native(o, size, mask=None):
    nat = _fb(o.rdram, size)
    nat_h = _hidden(o.hidden, size)
    for p in range(o.upscaling ** 2):
        up = _fb(o.upscaled_rdram[p], size)
        if mask is None:
            assert np.array_equal(up, nat), f"plane {p}"
            assert np.array_equal(_hidden(o.upscaled_hidden[p], size), nat_h), f"plane {p} hidden"
        else:
            assert np.array_equal(up[mask], nat[mask]), f"plane {p}"


def check_flat_rectangles(make, factor, size):
    """1-cycle texture rectangles (no native-resolution flag) with a flat
    primitive colour, pixel-aligned: every sample is covered exactly when
    its pixel is, so each plane holds the native image."""
    rng = np.random.default_rng(1000 + factor + size)
    cl = CommandList()
    _bind(cl, size=size)
    cl.set_other_modes(ONE_CYCLE)
    cl.set_combine(*PRIM)
    for _ in range(24):
        x0, y0 = int(rng.integers(0, W - 8)), int(rng.integers(0, H - 8))
        x1, y1 = int(rng.integers(x0 + 1, W)), int(rng.integers(y0 + 1, H))
        cl.set_prim_color(int(rng.integers(0, 1 << 32)))
        cl.texture_rectangle(0, x0 << 2, y0 << 2, x1 << 2, y1 << 2, 0, 0, 1 << 10, 1 << 10)
    cl.sync_full()
    o = make(factor)
    o.process(cl.words())
    o.flush()
    assert np.count_nonzero(_fb(o.rdram, size)) > W * H // 4
    _planes_equal_native(o, size)
    o.close()


def check_fill_rectangles(make, factor):
    """Fill rectangles carry the native-resolution flag: every sample shades
    as the pixel's first one, so the planes equal native, including the
    inclusive bottom line and the two-colour 16-bit fill pattern."""
    rng = np.random.default_rng(2000 + factor)
    cl = CommandList()
    _bind(cl)
    cl.set_other_modes(OtherModes(cycle_type=3))
    for _ in range(16):
        x0, y0 = int(rng.integers(0, W - 8)), int(rng.integers(0, H - 8))
        x1, y1 = int(rng.integers(x0, W - 1)), int(rng.integers(y0, H - 1))
        cl.set_fill_color(int(rng.integers(0, 1 << 32)))
        cl.fill_rectangle(x0 << 2, y0 << 2, x1 << 2, y1 << 2)
        cl.sync_pipe()
    cl.sync_full()
    o = make(factor)
    o.process(cl.words())
    o.flush()
    _planes_equal_native(o, rs.SIZ_16)
    o.close()


def check_copy_blit(make, factor):
    """A copy-mode texture rectangle at one texel per pixel: the samples of
    a pixel all read that pixel's texel (interpolate_st_copy drops the
    sub-pixel bits), so inside the rectangle every plane equals native."""
    rng = np.random.default_rng(3000 + factor)
    tex_w, tex_h = 32, 32
    tex_addr = 0x300000
    o = make(factor)
    texels = rng.integers(0, 1 << 16, size=tex_w * tex_h).astype(np.uint16)
    rs.rdram_write16(o.rdram, tex_addr, texels)
    cl = CommandList()
    _bind(cl)
    cl.set_other_modes(OtherModes(cycle_type=2))
    cl.set_texture_image(tex_addr, tex_w, rs.SIZ_16)
    cl.set_tile(0, rs.FMT_RGBA, rs.SIZ_16, 0, tex_w * 2)
    cl.sync_load()
    cl.load_tile(0, 0, 0, (tex_w - 1) << 2, (tex_h - 1) << 2)
    cl.sync_tile()
    cl.set_tile_size(0, 0, 0, (tex_w - 1) << 2, (tex_h - 1) << 2)
    x0, y0 = 100, 60
    x1, y1 = x0 + tex_w - 1, y0 + tex_h - 1
    cl.texture_rectangle(0, x0 << 2, y0 << 2, x1 << 2, y1 << 2, 0, 0, 4 << 10, 1 << 10)
    cl.sync_full()
    o.process(cl.words())
    o.flush()
    nat = _fb(o.rdram, rs.SIZ_16)
    assert np.array_equal(nat[y0:y1, x0:x1], texels.reshape(tex_h, tex_w)[:y1 - y0, :x1 - x0])
    inside = np.zeros((H, W), dtype=bool)
    inside[y0:y1, x0:x1] = True
    _planes_equal_native(o, rs.SIZ_16, mask=inside)
    # Nothing lands outside the rectangle's inclusive bounds.
    outside = np.ones((H, W), dtype=bool)
    outside[y0:y1 + 1, x0:x1 + 1] = False
    _planes_equal_native(o, rs.SIZ_16, mask=outside)
    o.close()


def _gradient_stream(rng, n):
    cl = CommandList()
    _bind(cl, size=rs.SIZ_32)
    cl.set_other_modes(ONE_CYCLE)
    cl.set_combine(*SHADE)
    xy = np.stack([rng.integers(0, W * 4, size=(n, 3)), rng.integers(0, H * 4, size=(n, 3))], axis=-1)
    rgba = rng.integers(0, 256, size=(n, 3, 4)).astype(np.float64)
    cl.extend_array(rs.triangle_commands(xy, rgba=rgba))
    cl.sync_full()
    return cl.words()


def check_gradient(make, factor):
    """Gouraud shading: sample (0, 0) of a pixel is evaluated where the
    native pixel is, so inside a triangle plane 0 agrees with native up to
    the rounding of the split gradients (2 of 255).  A gradient applied at
    native scale to S-times coordinates, or a span base left unscaled,
    drifts by tens to hundreds."""
    rng = np.random.default_rng(4000 + factor)
    o = make(factor)
    o.process(_gradient_stream(rng, 1))
    o.flush()
    nat = _fb(o.rdram, rs.SIZ_32).astype(np.int64)
    up0 = _fb(o.upscaled_rdram[0], rs.SIZ_32).astype(np.int64)
    covered = nat != 0
    inner = covered.copy()
    for dy in (-1, 0, 1):
        for dx in (-1, 0, 1):
            inner &= np.roll(np.roll(covered, dy, 0), dx, 1)
    inner &= up0 != 0
    assert inner.sum() > 500
    for shift in (24, 16, 8):
        d = np.abs(((nat >> shift) & 0xFF) - ((up0 >> shift) & 0xFF))[inner]
        assert d.max() <= 2, (shift, d.max())
    o.close()


def check_host_writes(make):
    """The pre update broadcasts native memory that changed since the domain
    last agreed with it, over the framebuffer a pass draws (width x deduced
    height); memory outside it stays as it was in the planes."""
    rng = np.random.default_rng(5)
    f = 2
    o = make(f)
    host = rng.integers(1, 1 << 16, size=W * H).astype(np.uint16)
    rs.rdram_write16(o.rdram, COLOR, host)
    far = 0x500000
    rs.rdram_write16(o.rdram, far, host[:1024])
    o.hidden[COLOR // 2: COLOR // 2 + W * H] = rng.integers(0, 4, size=W * H)
    cl = CommandList()
    _bind(cl)
    cl.set_other_modes(ONE_CYCLE)
    cl.set_combine(*PRIM)
    cl.set_prim_color(0xFF0000FF)
    cl.texture_rectangle(0, 10 << 2, (H - 2) << 2, 20 << 2, (H - 1) << 2, 0, 0, 1 << 10, 1 << 10)
    cl.sync_full()
    o.process(cl.words())
    o.flush()
    nat = _fb(o.rdram, rs.SIZ_16)
    assert np.array_equal(nat.ravel()[:W], host[:W])
    # The pass's deduced height ends at its last drawn line, H - 2.
    region = np.zeros((H, W), dtype=bool)
    region[:H - 1] = True
    _planes_equal_native(o, rs.SIZ_16, mask=region)
    for p in range(f * f):
        assert np.array_equal(_hidden(o.upscaled_hidden[p], rs.SIZ_16)[:W * (H - 1)],
                              _hidden(o.hidden, rs.SIZ_16)[:W * (H - 1)])
        assert not np.any(_fb(o.upscaled_rdram[p], rs.SIZ_16)[H - 1])
        assert not np.any(rs.rdram_read16(o.upscaled_rdram[p], far, 1024))
    o.close()


def check_post_update(make):
    """After an upscaled pass the domain's reference copy equals native
    memory, so a later pass over the same framebuffer leaves the S-times
    result of the first in the planes instead of broadcasting native
    pixels over it."""
    f = 2
    rng = np.random.default_rng(6)
    o = make(f)
    o.process(_gradient_stream(rng, 3))
    o.flush()
    before = o.upscaled_rdram.copy()
    assert any(not np.array_equal(_fb(before[p], rs.SIZ_32), _fb(o.rdram, rs.SIZ_32)) for p in range(f * f))
    cl = CommandList()
    _bind(cl, size=rs.SIZ_32)
    cl.set_other_modes(ONE_CYCLE)
    cl.set_combine(*PRIM)
    cl.set_prim_color(0x12345678)
    cl.texture_rectangle(0, 4 << 2, (H - 4) << 2, 12 << 2, (H - 1) << 2, 0, 0, 1 << 10, 1 << 10)
    cl.sync_full()
    o.process(cl.words())
    o.flush()
    untouched = np.ones((H, W), dtype=bool)
    untouched[H - 4:H - 1, 4:12] = False
    for p in range(f * f):
        assert np.array_equal(_fb(o.upscaled_rdram[p], rs.SIZ_32)[untouched], _fb(before[p], rs.SIZ_32)[untouched])
        rect = _fb(o.upscaled_rdram[p], rs.SIZ_32)[H - 4:H - 1, 4:12]
        assert np.array_equal(rect, _fb(o.rdram, rs.SIZ_32)[H - 4:H - 1, 4:12])
        assert np.all(rect >> 8 == 0x123456)
    o.close()


def check_native_only(make, bind):
    """Framebuffers narrower than 256 pixels or not RGBA5551/RGBA8888 render
    natively only: their output reaches the planes only through the pre
    update of a later upscaled pass over the same memory."""
    f = 2
    o = make(f)
    width = bind.get("width", W)
    cl = CommandList()
    _bind(cl, **bind)
    cl.set_other_modes(ONE_CYCLE)
    cl.set_combine(*PRIM)
    cl.set_prim_color(0xF0F0F0FF)
    cl.texture_rectangle(0, 0, 0, (width // 2) << 2, 40 << 2, 0, 0, 1 << 10, 1 << 10)
    cl.sync_full()
    o.process(cl.words())
    o.flush()
    assert np.count_nonzero(o.rdram[COLOR // 4: COLOR // 4 + W * 10])
    assert not np.any(o.upscaled_rdram)
    # An upscaled pass over the same memory pulls the native output in.
    cl = CommandList()
    _bind(cl)
    cl.set_other_modes(ONE_CYCLE)
    cl.set_combine(*PRIM)
    cl.set_prim_color(0x0F0F0F0F)
    cl.texture_rectangle(0, 0, (H - 2) << 2, 4 << 2, (H - 1) << 2, 0, 0, 1 << 10, 1 << 10)
    cl.sync_full()
    o.process(cl.words())
    o.flush()
    _planes_equal_native(o, rs.SIZ_16)
    o.close()


def check_texture_columns(make, factor):
    """Plane order and the texture-coordinate gradient: a 1-cycle texture
    rectangle at two texels per pixel (point sampled) over a texture whose
    texels vary along s only.  Sample X = S x + sx of the S-times image
    reads texel s = 2 X / S, so plane sy * S + sx at pixel x holds texel
    2 x + (2 sx) // S, whatever sy."""
    f = factor
    tex_w, tex_h = 64, 16
    tex_addr = 0x300000
    o = make(f)
    col = (np.arange(tex_w, dtype=np.uint16) * 0x0421 * 3 + 0x8000) | 1
    rs.rdram_write16(o.rdram, tex_addr, np.tile(col, tex_h))
    cl = CommandList()
    _bind(cl)
    # Point sampling (sample type 0) with the filter in bilerp rather than
    # YUV-convert mode, so texels reach the combiner unchanged.
    cl.set_other_modes(OtherModes(cycle_type=0, bilerp0=True, bilerp1=True, blend=((0, 0, 1, 0), (0, 0, 1, 0))))
    cl.set_combine(*TEX0)
    cl.set_texture_image(tex_addr, tex_w, rs.SIZ_16)
    cl.set_tile(0, rs.FMT_RGBA, rs.SIZ_16, 0, tex_w * 2)
    cl.sync_load()
    cl.load_tile(0, 0, 0, (tex_w - 1) << 2, (tex_h - 1) << 2)
    cl.sync_tile()
    cl.set_tile_size(0, 0, 0, (tex_w - 1) << 2, (tex_h - 1) << 2)
    x0, y0, n = 40, 30, tex_w // 2
    cl.texture_rectangle(0, x0 << 2, y0 << 2, (x0 + n) << 2, (y0 + 8) << 2, 0, 0, 2 << 10, 1 << 10)
    cl.sync_full()
    o.process(cl.words())
    o.flush()
    k = np.arange(n)
    nat = _fb(o.rdram, rs.SIZ_16)[y0:y0 + 8, x0:x0 + n]
    assert np.array_equal(nat >> 1, np.broadcast_to(col[2 * k] >> 1, nat.shape))
    for sy in range(f):
        for sx in range(f):
            up = _fb(o.upscaled_rdram[sy * f + sx], rs.SIZ_16)[y0:y0 + 8, x0:x0 + n]
            exp = col[2 * k + (2 * sx) // f] >> 1
            assert np.array_equal(up >> 1, np.broadcast_to(exp, up.shape)), (sx, sy)
    o.close()


def check_vi_planes(make, factor):
    """Closed form of the upscaled fetch (extract_vram.comp:61-70): with
    unit scale registers and no filters, output pixel (X, Y) is sample
    plane (Y mod f) * f + (X mod f) at native pixel (X / f, Y / f).  The
    planes are written directly and native RDRAM equals the domain's
    reference copy (both zero), so the scanout's domain update leaves them
    alone."""
    f = factor
    o = make(f)
    tvi.default_vi(o, fmt=tvi.RGBA8888, aa=tvi.AA_RESAMP_ONLY)
    w, h = 640, 240
    ny, nx = np.mgrid[0:h, 0:w].astype(np.uint32)
    for p in range(f * f):
        words = ((nx & 0xFF) << 24) | ((ny & 0xFF) << 16) | (np.uint32(p * 13 & 0xFF) << 8)
        o.upscaled_rdram[p, 4096 // 4: 4096 // 4 + w * h] = words.ravel()
    o.set_vi_register(rs.VI_ORIGIN, 4096)
    o.set_vi_register(rs.VI_WIDTH, w)
    o.set_vi_register(rs.VI_V_START, tvi.start_reg(tvi.V_OFF, tvi.V_OFF + h * 2))
    img = o.scanout(tvi.suite_options())
    assert img.shape == (h * f, w * f)
    Y, X = np.mgrid[0:h * f, 0:w * f]
    exp_r, exp_g = (X // f) & 0xFF, (Y // f) & 0xFF
    exp_b = ((Y % f) * f + (X % f)) * 13 & 0xFF
    vis = (X >= 8 * f) & (X < (640 - 7) * f)
    got = img.astype(np.int64)
    assert np.array_equal(got & 0xFF, np.where(vis, exp_r, 0))
    assert np.array_equal((got >> 8) & 0xFF, np.where(vis, exp_g, 0))
    assert np.array_equal((got >> 16) & 0xFF, np.where(vis, exp_b, 0))
    o.close()


# ---------------------------------------------------------------------------
# Super-sampled readback: passes render at scale only and resolve into
# native RDRAM (update_upscaled_domain_resolve.comp).
# ---------------------------------------------------------------------------
def _ssaa(make, factor, dither=False):
    o = make(factor)
    o.enable_super_sampled_readback(dither)
    return o


def _mean_check(native, planes, dither):
    """native (H, W) RGBA5551 words against the rounded mean of the plane
    samples (N, H, W) per 5-bit channel and the alpha bit; with dither the
    4x4 ordered offsets b / 8 (b = 0..7 from the Bayer-style table) are
    added before the floor."""
    n = planes.shape[0]
    Y, X = np.mgrid[0:native.shape[0], 0:native.shape[1]]
    bayer = np.array([0, 4, 1, 5, 4, 0, 5, 1, 3, 7, 2, 6, 7, 3, 6, 2])[(Y & 3) * 4 + (X & 3)]
    for shift, bits in ((11, 5), (6, 5), (1, 5), (0, 1)):
        mean = ((planes.astype(np.int64) >> shift) & ((1 << bits) - 1)).sum(axis=0) / n
        exp = np.floor(mean + (bayer / 8.0 if dither else 0.5)).astype(np.int64)
        got = (native.astype(np.int64) >> shift) & ((1 << bits) - 1)
        assert np.array_equal(got, exp), (shift, int(np.count_nonzero(got != exp)))


def check_ssaa_flat(make, factor, size):
    """Pixel-aligned flat rectangles: all samples of a pixel agree, so the
    resolved native colour is exactly the native rendering's; hidden RDRAM
    is not written by the resolve."""
    rng = np.random.default_rng(9000 + factor + size)
    cl = CommandList()
    _bind(cl, size=size)
    cl.set_other_modes(ONE_CYCLE)
    cl.set_combine(*PRIM)
    for _ in range(16):
        x0, y0 = int(rng.integers(0, W - 8)), int(rng.integers(0, H - 8))
        x1, y1 = int(rng.integers(x0 + 1, W)), int(rng.integers(y0 + 1, H))
        cl.set_prim_color(int(rng.integers(0, 1 << 32)))
        cl.texture_rectangle(0, x0 << 2, y0 << 2, x1 << 2, y1 << 2, 0, 0, 1 << 10, 1 << 10)
    cl.sync_full()
    o = _ssaa(make, factor)
    o.process(cl.words())
    o.flush()
    ref = make(1)
    ref.process(cl.words())
    ref.flush()
    assert np.count_nonzero(_fb(o.rdram, size)) > W * H // 4
    assert np.array_equal(_fb(o.rdram, size), _fb(ref.rdram, size))
    assert np.all(o.hidden == 3)
    o.close()
    ref.close()


def check_ssaa_mean(make, factor, dither):
    """Gouraud triangles into RGBA5551: after the pass every native pixel of
    the framebuffer is the (dithered) mean of its samples.  Pixels no sample
    wrote hold what the pre update copied to every plane, so the mean is
    their own value."""
    rng = np.random.default_rng(9100 + factor + dither)
    cl = CommandList()
    _bind(cl)
    cl.set_other_modes(ONE_CYCLE)
    cl.set_combine(*SHADE)
    xy = np.stack([rng.integers(0, W * 4, size=(6, 3)), rng.integers(0, H * 4, size=(6, 3))], axis=-1)
    rgba = rng.integers(0, 256, size=(6, 3, 4)).astype(np.float64)
    cl.extend_array(rs.triangle_commands(xy, rgba=rgba))
    cl.sync_full()
    o = _ssaa(make, factor, dither)
    o.process(cl.words())
    o.flush()
    nat = _fb(o.rdram, rs.SIZ_16)
    planes = np.stack([_fb(o.upscaled_rdram[p], rs.SIZ_16) for p in range(factor * factor)])
    assert len(np.unique(nat)) > 64
    _mean_check(nat, planes, dither)
    o.close()


def check_ssaa_depth(make, factor):
    """Z-buffered triangles: the resolved depth is the first sample's."""
    rng = np.random.default_rng(9200 + factor)
    cl = CommandList()
    # Clear Z through a fill pass (at scale too), then draw.
    _bind(cl, addr=rs.DEPTH_ADDR)
    cl.set_other_modes(OtherModes(cycle_type=3))
    cl.set_fill_color(0xFFFCFFFC)
    cl.fill_rectangle(0, 0, (W - 1) << 2, (H - 1) << 2)
    cl.sync_pipe()
    _bind(cl)
    cl.set_other_modes(OtherModes(cycle_type=0, depth_test=True, depth_write=True, blend=((0, 0, 1, 0), (0, 0, 1, 0))))
    cl.set_combine(*SHADE)
    xy = np.stack([rng.integers(0, W * 4, size=(8, 3)), rng.integers(0, H * 4, size=(8, 3))], axis=-1)
    rgba = rng.integers(0, 256, size=(8, 3, 4)).astype(np.float64)
    z = rng.integers(0, 0x40000, size=(8, 3)).astype(np.float64)
    cl.extend_array(rs.triangle_commands(xy, rgba=rgba, z=z))
    cl.sync_full()
    o = _ssaa(make, factor)
    o.process(cl.words())
    o.flush()
    nat = _fb(o.rdram, rs.SIZ_16, addr=rs.DEPTH_ADDR)
    assert np.count_nonzero(nat != 0xFFFC) > 1000
    assert np.array_equal(nat, _fb(o.upscaled_rdram[0], rs.SIZ_16, addr=rs.DEPTH_ADDR))
    o.close()


def check_ssaa_host_writes(make):
    """Memory the pass does not draw keeps what the host wrote."""
    rng = np.random.default_rng(9300)
    o = _ssaa(make, 2)
    host = rng.integers(1, 1 << 16, size=W * H).astype(np.uint16)
    rs.rdram_write16(o.rdram, COLOR, host)
    cl = CommandList()
    _bind(cl)
    cl.set_other_modes(ONE_CYCLE)
    cl.set_combine(*PRIM)
    cl.set_prim_color(0xFF00FFFF)
    cl.texture_rectangle(0, 40 << 2, 40 << 2, 60 << 2, 50 << 2, 0, 0, 1 << 10, 1 << 10)
    cl.sync_full()
    o.process(cl.words())
    o.flush()
    nat = _fb(o.rdram, rs.SIZ_16)
    keep = np.ones((H, W), dtype=bool)
    keep[40:50, 40:60] = False
    assert np.array_equal(nat[keep], host.reshape(H, W)[keep])
    assert np.all(nat[40:50, 40:60] >> 1 == 0xF83E >> 1)  # magenta in RGBA5551
    o.close()


# ---------------------------------------------------------------------------
# The pins, and a planted mistake for each that its pin must catch.
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("factor", [2, 4, 8])
@pytest.mark.parametrize("size", [rs.SIZ_16, rs.SIZ_32])
def test_flat_rectangles_replicate_into_every_plane(oracle_mod, factor, size):
    check_flat_rectangles(_maker(oracle_mod), factor, size)


@pytest.mark.parametrize("factor", [2, 4])
def test_fill_rectangles_replicate(oracle_mod, factor):
    check_fill_rectangles(_maker(oracle_mod), factor)


@pytest.mark.parametrize("factor", [2, 4])
def test_copy_blit_replicates_texels(oracle_mod, factor):
    check_copy_blit(_maker(oracle_mod), factor)


@pytest.mark.parametrize("factor", [2, 4])
def test_gradient_first_sample_matches_native(oracle_mod, factor):
    check_gradient(_maker(oracle_mod), factor)


@pytest.mark.parametrize("factor", [2, 4])
def test_texture_columns_by_plane(oracle_mod, factor):
    check_texture_columns(_maker(oracle_mod), factor)


def test_host_writes_reach_planes_through_a_drawn_framebuffer(oracle_mod):
    check_host_writes(_maker(oracle_mod))


def test_post_update_keeps_upscaled_content(oracle_mod):
    check_post_update(_maker(oracle_mod))


@pytest.mark.parametrize("bind", [dict(width=255), dict(size=rs.SIZ_8, fmt=rs.FMT_I),
                                  dict(size=rs.SIZ_16, fmt=rs.FMT_IA)])
def test_native_only_passes(oracle_mod, bind):
    check_native_only(_maker(oracle_mod), bind)


@pytest.mark.parametrize("factor", [2, 4])
def test_upscaled_scanout_reads_sample_planes(oracle_mod, factor):
    check_vi_planes(_maker(oracle_mod), factor)


@pytest.mark.parametrize("factor", [2, 4])
@pytest.mark.parametrize("size", [rs.SIZ_16, rs.SIZ_32])
def test_ssaa_flat_rectangles_resolve_to_native(oracle_mod, factor, size):
    check_ssaa_flat(_maker(oracle_mod), factor, size)


@pytest.mark.parametrize("factor,dither", [(2, False), (2, True), (4, False), (4, True)])
def test_ssaa_resolve_is_the_sample_mean(oracle_mod, factor, dither):
    check_ssaa_mean(_maker(oracle_mod), factor, dither)


@pytest.mark.parametrize("factor", [2, 4])
def test_ssaa_depth_is_the_first_sample(oracle_mod, factor):
    check_ssaa_depth(_maker(oracle_mod), factor)


def test_ssaa_keeps_undrawn_host_memory(oracle_mod):
    check_ssaa_host_writes(_maker(oracle_mod))


PLANTS = {
    30: lambda m: check_gradient(m, 2),              # colour x gradient not split by the factor
    31: lambda m: check_flat_rectangles(m, 2, rs.SIZ_16),  # H edge left at native scale
    32: lambda m: check_texture_columns(m, 2),       # sample planes transposed
    33: check_post_update,                           # post update skipped
    34: check_host_writes,                           # pre update leaves the hidden bits
    35: lambda m: check_copy_blit(m, 2),             # copy-mode texel steps at S-times rate
    36: lambda m: check_vi_planes(m, 2),             # VI fetch reads transposed planes
    37: lambda m: check_texture_columns(m, 2),       # s/t x gradient not split by the factor
    38: lambda m: check_ssaa_mean(m, 2, False),      # resolve mean truncated instead of rounded
    39: lambda m: check_ssaa_mean(m, 2, True),       # dither offsets transposed
    40: lambda m: check_ssaa_depth(m, 2),            # resolved depth from the second sample
}


def _planted(n):
    out_dir = os.path.join(ORACLE, "build", "planted")
    os.makedirs(out_dir, exist_ok=True)
    out = os.path.join(out_dir, f"librdp_oracle_plant{n}.so")
    srcs = [os.path.join(ORACLE, f) for f in ("rdp_oracle.c", "vi_oracle.c", "rdp_oracle.h", "vi_oracle.h")]
    if not os.path.exists(out) or os.path.getmtime(out) < max(os.path.getmtime(s) for s in srcs):
        subprocess.run(["gcc", "-O1", "-fwrapv", "-fPIC", "-std=c99", "-shared", f"-DRDPO_PLANT={n}", "-o", out,
                        srcs[0], srcs[1], "-lm"], check=True)
    return out


@pytest.mark.parametrize("plant", sorted(PLANTS))
def test_planted_mistake_fails_its_pin(oracle_mod, plant):
    make = _maker(oracle_mod, _planted(plant))
    with pytest.raises(AssertionError):
        PLANTS[plant](make)


def test_factor_validation(oracle_mod):
    o = oracle_mod.Oracle(rs.RDRAM_SIZE)
    for bad in (0, 1, 3, 16):
        with pytest.raises(ValueError):
            o.enable_upscaling(bad)
    o.enable_upscaling(2)
    with pytest.raises(ValueError):
        o.enable_upscaling(4)
    o.close()
