"""Pins of the oracle's rendering functions by what the mathematics or the
RDP's documented formats fix, written independently of the oracle's own
formulas, and for each pin a planted mistake (the oracle built with
-DRDPO_PLANT=n, see rdp_oracle.c) that the pin must catch.

What each pin states:
  span walker   a sample is covered exactly when it lies between the exact
                H and M/L edge positions of its sub-scanline (27-bit wrap),
                inside the scissor and the y range, on a kept interlace line
                (span_setup.comp:87-227 with coverage.h:31-44);
  interpolation the direct forms of interpolation.h:144-253 agree with the
                span-based path; on an integer plane the shade, z and st
                values equal the plane at the first covered sample (floored);
  texel decode  every texel value of every format, placed by the TMEM layout
                (odd lines swap 32-bit words), decodes as the format defines
                (bit replication for 5/4/3-bit fields, TLUT banks, YUV);
  addressing    wrap is s mod 2^mask, mirror reflects odd periods, clamp
                holds the tile's edge texel, over every coordinate;
  filtering     a constant texture stays constant; a linear ramp gives the
                exact interpolant rounded half up, for 3-tap and 4-tap;
  LOD           power-of-two texel steps select their level with fraction 0;
                in between, the fraction is linear in the step;
  combiner      every selector reads the input the RDP's tables name;
  depth test    one invariant per Z mode (opaque monotone, transparent
                front-only, decal window symmetric, interpenetrate = opaque
                off the overlap case);
  blender       alpha / 1-alpha weights sum to one (out = weighted mean);
                the final-cycle divider agrees; the early outs hold;
  coverage      clamp, wrap, zap and save destination modes;
  framebuffer   pack/unpack round trips of all five colour formats and the
                depth format, hidden bits included; fill by address parity;
                copy as identity.
"""
import ctypes
import math
import os
import subprocess
from fractions import Fraction

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE = os.path.join(ROOT, "oracle")

TRI_FLIP, TRI_DO_OFFSET, TRI_SKIP_XFRAC, TRI_INTERLACE, TRI_KEEP_ODD = 1, 2, 4, 8, 16
TF_RGBA, TF_YUV, TF_CI, TF_IA, TF_I = 0, 1, 2, 3, 4
FB_I4, FB_I8, FB_RGBA5551, FB_IA88, FB_RGBA8888 = 0, 1, 2, 3, 4
TILE_CLAMP_S, TILE_MIRROR_S = 1, 2

vp = ctypes.c_void_p
ci = ctypes.c_int
u32 = ctypes.c_uint32
i32 = ctypes.c_int32


def _bind(L):
    L.rdpo_create.restype = vp
    L.rdpo_create.argtypes = [u32]
    L.rdpo_destroy.argtypes = [vp]
    for n in ("rdpo_tmem", "rdpo_rdram", "rdpo_hidden_rdram"):
        getattr(L, n).restype = vp
        getattr(L, n).argtypes = [vp]
    L.rdpo_test_span_setup.argtypes = [vp, vp, vp, ci, vp]
    L.rdpo_test_coverage.restype = ci
    L.rdpo_test_coverage.argtypes = [vp, vp, ci]
    L.rdpo_test_interpolate_rgba.argtypes = [vp, vp, ci, ci, vp]
    L.rdpo_test_interpolate_stz.argtypes = [vp, vp, ci, ci, ci, ci, ci, vp]
    L.rdpo_test_fetch_texel.argtypes = [vp, vp, u32, u32, ci, ci, u32, u32, vp]
    L.rdpo_test_fetch_yuv.argtypes = [vp, vp, u32, u32, u32, vp]
    L.rdpo_test_texel_mask.restype = ci
    L.rdpo_test_texel_mask.argtypes = [vp, ci, ci]
    L.rdpo_test_sample_texture.argtypes = [vp, vp, vp, ci, ci, ci, ci, ci, ci, vp, vp, vp]
    L.rdpo_test_compute_lod.argtypes = [u32, u32, ci, vp, ci, ci, ci, ci, vp]
    L.rdpo_test_select.argtypes = [vp, vp, ci, ci, ci, vp]
    L.rdpo_test_depth_test.restype = ci
    L.rdpo_test_depth_test.argtypes = [i32, i32, i32, u32, u32, ci, ci, ci, ci, ci, ci, vp]
    L.rdpo_test_blender.argtypes = [vp, vp, vp, ci, ci, ci, ci, ci, ci, ci, vp]
    L.rdpo_test_blend_coverage.restype = ci
    L.rdpo_test_blend_coverage.argtypes = [ci, ci, ci, ci]
    L.rdpo_test_load_color.argtypes = [vp, u32, u32, vp]
    L.rdpo_test_store_color.argtypes = [vp, u32, u32, vp]
    L.rdpo_test_load_depth.argtypes = [vp, u32, vp]
    L.rdpo_test_store_depth.argtypes = [vp, u32, u32, u32]
    L.rdpo_test_fill_color.argtypes = [u32, u32, ci, u32, vp]
    L.rdpo_test_copy_pipeline.argtypes = [u32, ci, u32, vp]
    L.rdpo_test_z_decompress.restype = i32
    L.rdpo_test_z_decompress.argtypes = [u32]
    return L


class Rdp:
    """An oracle instance (1 MiB RDRAM) with numpy views of its memories."""

    def __init__(self, L):
        self.L = L
        self.h = L.rdpo_create(1 << 20)
        self.tmem = np.frombuffer((ctypes.c_uint8 * 4096).from_address(L.rdpo_tmem(self.h)), np.uint8)
        self.rdram = np.frombuffer((ctypes.c_uint8 * (1 << 20)).from_address(L.rdpo_rdram(self.h)), np.uint8)
        self.hidden = np.frombuffer((ctypes.c_uint8 * (1 << 19)).from_address(L.rdpo_hidden_rdram(self.h)), np.uint8)

    def close(self):
        self.L.rdpo_destroy(self.h)


def _i32(v):
    return ((int(v) + (1 << 31)) % (1 << 32)) - (1 << 31)


def _sext(v, bits):
    v = int(v) & ((1 << bits) - 1)
    return v - (1 << bits) if v >> (bits - 1) else v


def _clamp9(c):
    return min(max(_sext(c - 0x80, 9) + 0x80, 0), 255)


def _arr(values, dtype=np.int32):
    return np.ascontiguousarray(np.array(values, dtype=np.int64).astype(dtype))


# ---------------------------------------------------------------------------
# Span walker: exact edges.
# ---------------------------------------------------------------------------
SAMPLE_X8 = (0, 4, 2, 6)  # eighths of a pixel: lines 0/2 at 0 and 4, lines 1/3 at 2 and 6


def _model_mask(tri, sc, x, y):
    """Coverage of pixel (x, y) from the exact edge positions.  X values carry
    15 fraction bits, Y values 2; a sub-scanline's H and M edges start from the
    whole scanline of YH, its L edge from YM."""
    xh, xm, xl, yh, ym, yl, dxh, dxm, dxl, flags = tri
    if flags & TRI_INTERLACE and (y & 1) != (1 if flags & TRI_KEEP_ODD else 0):
        return 0
    ylo, yhi = max(yh, sc[1]), min(yl, sc[3])
    top = yh - (yh % 4)
    mask = 0
    for k in range(8):
        line = k >> 1
        ys = 4 * y + line
        if ys < ylo or ys >= yhi:
            continue
        h = _sext(xh + (ys - top) * dxh, 27)
        minor = _sext(xm + (ys - top) * dxm, 27) if ys < ym else _sext(xl + (ys - ym) * dxl, 27)
        left, right = (h, minor) if flags & TRI_FLIP else (minor, h)
        xs = 8 * x + SAMPLE_X8[((line & 1) << 1) | (k & 1)]
        # Sample at xs / 8 px; edges at value / 2^15 px; scissor in quarter pixels.
        if left <= xs * 4096 < right and 2 * sc[0] <= xs < 2 * sc[2]:
            mask |= 1 << k
    return mask


def _random_tri(rng, W, H, big=False):
    yh = int(rng.integers(-6, 4 * H - 4))
    ym = yh + int(rng.integers(0, 2 * H))
    yl = ym + int(rng.integers(0, 2 * H))
    xh = int(rng.integers(-4 << 15, (W + 4) << 15))
    xm = xh + int(rng.integers(-(W << 14), W << 14))
    xl = xh + int(rng.integers(-(W << 14), W << 14))
    slopes = [int(rng.integers(-(1 << 15), 1 << 15)) for _ in range(3)]
    flags = (TRI_FLIP if rng.random() < 0.5 else 0)
    if rng.random() < 0.2:
        flags |= TRI_INTERLACE | (TRI_KEEP_ODD if rng.random() < 0.5 else 0)
    if big:  # 27-bit wrap of the edge values
        xh += (1 << 27) * int(rng.choice([-1, 1]))
    return [xh, xm, xl, yh, ym, yl] + slopes + [flags]


def _span(L, tri, attr, sc, y):
    out = np.zeros(21, np.int32)
    L.rdpo_test_span_setup(_arr(tri).ctypes.data, _arr(attr).ctypes.data, _arr(sc, np.uint32).ctypes.data, y,
                           out.ctypes.data)
    return out


def pin_span_walker(L):
    rng = np.random.default_rng(101)
    W, H = 40, 24
    attr = [0] * 32
    for n in range(160):
        tri = _random_tri(rng, W, H, big=(n % 16 == 0))
        xlo = int(rng.integers(0, 4 * 8))
        ylo = int(rng.integers(0, 4 * 6))
        sc = [xlo, ylo, xlo + int(rng.integers(4, 4 * W)), ylo + int(rng.integers(4, 4 * H))]
        for y in range(max(tri[3] // 4 - 1, 0), min(tri[5] // 4 + 2, H + 2)):
            sp = _span(L, tri, attr, sc, y)
            xl = _arr(sp[8:12], np.uint16)
            xr = _arr(sp[12:16], np.uint16)
            for x in range(0, W + 2):
                got = L.rdpo_test_coverage(xl.ctypes.data, xr.ctypes.data, x) if sp[20] else 0
                want = _model_mask(tri, sc, x, y)
                assert got == want, (n, tri, sc, x, y, hex(got), hex(want))


# ---------------------------------------------------------------------------
# Attribute interpolation.
# ---------------------------------------------------------------------------
def _direct_rgba(tri, attr, x, y, coverage, ch):
    """interpolation.h:144-194 for one channel.  The X fields here carry 15
    fraction bits (the decoder halves X), so the pixel and the 8-bit x
    fraction sit one bit lower than in that listing."""
    xh, dxh, flags = tri[0], tri[6], tri[9]
    A, dx, de, dy = attr[ch], attr[4 + ch], attr[8 + ch], attr[12 + ch]
    yb = tri[3] >> 2
    xline = _i32(xh + (y - yb) * _i32(dxh << 2))
    diff = 0
    if flags & TRI_DO_OFFSET:
        xline = _i32(xline + 3 * dxh)
        deh, dyh = de & ~0x1FF, dy & ~0x1FF
        diff = deh - (deh >> 2) - dyh + (dyh >> 2)
    base_x = xline >> 15
    xfrac = 0 if flags & TRI_SKIP_XFRAC else (xline >> 7) & 0xFF
    v = _i32(A + _i32(de * (y - yb)))
    v = _i32(_i32((v & ~0x1FF) + diff - _i32(xfrac * ((dx >> 8) & ~1)))) & ~0x3FF
    v = _i32(v + _i32((dx & ~0x1F) * (x - base_x)))
    first = (coverage & -coverage).bit_length() - 1
    yoff = first >> 1
    xoff = ((first & 1) << 1) + (yoff & 1)
    s = (v >> 14) << 2
    s += xoff * (dx >> 14) + yoff * (dy >> 14)
    return _clamp9(s >> 4)


def _interp_rgba(L, sp, attr, dx, cov):
    out = np.zeros(4, np.int32)
    L.rdpo_test_interpolate_rgba(_arr(sp[0:4]).ctypes.data, _arr(attr).ctypes.data, dx, cov, out.ctypes.data)
    return out


def _interp_stz(L, sp, attr, dx, cov, persp=0):
    out = np.zeros(8, np.int32)
    L.rdpo_test_interpolate_stz(_arr(sp[4:8]).ctypes.data, _arr(attr).ctypes.data, dx, cov, persp, 0, 1,
                                out.ctypes.data)
    return out


def pin_interpolation_direct_form(L):
    rng = np.random.default_rng(202)
    sc = [0, 0, 4095, 4095]
    for n in range(300):
        tri = _random_tri(rng, 64, 48)
        tri[9] = (tri[9] & TRI_FLIP) | (TRI_DO_OFFSET if rng.random() < 0.5 else 0) | \
            (TRI_SKIP_XFRAC if rng.random() < 0.2 else 0)
        attr = [int(v) for v in rng.integers(-(1 << 31), 1 << 31, size=32, dtype=np.int64)]
        for _ in range(4):
            y = int(rng.integers(0, 48))
            sp = _span(L, tri, attr, sc, y)
            for _ in range(6):
                x = int(rng.integers(-8, 72))
                cov = int(rng.integers(1, 256))
                got = _interp_rgba(L, sp, attr, x - int(sp[16]), cov)
                want = [_direct_rgba(tri, attr, x, y, cov, ch) for ch in range(4)]
                assert list(got) == want, (n, x, y, cov, list(got), want)


def pin_interpolation_plane(L):
    """Integer planes: shade, z and st at a covered pixel are the plane's
    value at the first covered sample (x + xoff/4, y + yoff/4), floored; z
    keeps 3 fraction bits."""
    rng = np.random.default_rng(303)
    sc = [0, 0, 4095, 4095]
    xoff4 = (0, 2, 1, 3, 0, 2, 1, 3)  # sample k's x offset in quarter pixels
    for n in range(300):
        X0, Y0 = int(rng.integers(0, 32)), int(rng.integers(0, 32))
        tri = [X0 << 15, X0 << 15, (X0 + 40) << 15, 4 * Y0, 4 * Y0 + 40, 4 * Y0 + 80, 0, 1 << 13, 0, TRI_FLIP]
        c = [int(v) for v in rng.integers(0, 256, 4)]
        gx = [int(v) for v in rng.integers(-3, 4, 4)]
        gy = [int(v) for v in rng.integers(-3, 4, 4)]
        zc, zx, zy = int(rng.integers(0x400, 0x7800)), int(rng.integers(-40, 40)), int(rng.integers(-40, 40))
        sc_, sx, sy = int(rng.integers(-500, 500)), int(rng.integers(-9, 9)), int(rng.integers(-9, 9))
        attr = [v << 16 for v in c] + [v << 16 for v in gx] + [v << 16 for v in gy] + [v << 16 for v in gy]
        attr += [sc_ << 16, 0, zc << 16, 0, sx << 16, 0, zx << 16, 0, sy << 16, 0, zy << 16, 0, sy << 16, 0, zy << 16, 0]
        for _ in range(8):
            y = Y0 + int(rng.integers(0, 18))
            x = X0 + int(rng.integers(0, 30))
            cov = int(rng.integers(1, 256))
            k = (cov & -cov).bit_length() - 1
            qx, qy = Fraction(xoff4[k], 4), Fraction(k >> 1, 4)
            sp = _span(L, tri, attr, sc, y)
            dx = x - int(sp[16])
            got = _interp_rgba(L, sp, attr, dx, cov)
            for ch in range(4):
                value = c[ch] + gx[ch] * (x - X0 + qx) + gy[ch] * (y - Y0 + qy)
                assert got[ch] == _clamp9(math.floor(value)), (n, ch, x, y, k, got[ch], value)
            stz = _interp_stz(L, sp, attr, dx, cov)
            z = 8 * (zc + zx * (x - X0 + qx) + zy * (y - Y0 + qy))
            assert stz[6] == min(max(math.floor(z), 0), 0x3FFFF), (n, x, y, k, stz[6], z)
            assert stz[0] == sc_ + sx * (x - X0) + sy * (y - Y0)


# ---------------------------------------------------------------------------
# TMEM layout and texel formats.
# ---------------------------------------------------------------------------
def _tile(fmt, size, offset=0, stride=64, palette=0, mask_s=0, mask_t=0, shift_s=0, shift_t=0, slo=0, shi=0,
          tlo=0, thi=0, flags=0):
    return _arr([slo, shi, tlo, thi, offset, stride, fmt, size, palette, mask_s, shift_s, mask_t, shift_t, flags],
                np.uint32)


def _texel_byte(offset, stride, s, t, bits, wrap=0xFFF):
    """N64 TMEM byte address of texel (s, t): lines of `stride` bytes; on odd
    lines the two 32-bit halves of every 64-bit word trade places."""
    b = (offset + stride * t + (s * bits) // 8) & wrap
    return b ^ 4 if t & 1 else b


def _put8(tm, b, v):
    tm[(b & 0xFFF) ^ 3] = v


def _put16(tm, b, v):
    _put8(tm, b, (v >> 8) & 0xFF)
    _put8(tm, b + 1, v & 0xFF)


def _ex5(c):
    return (c << 3) | (c >> 2)


def _rgba16(w):
    return (_ex5((w >> 11) & 31), _ex5((w >> 6) & 31), _ex5((w >> 1) & 31), 255 if w & 1 else 0)


def _ia16(w):
    return (w >> 8, w >> 8, w >> 8, w & 0xFF)


def _fetch(L, r, tile, s, t, tlut=0, tlut_type=0, lut_offset=0, addr_xor=0):
    out = np.zeros(4, np.int32)
    L.rdpo_test_fetch_texel(r.h, tile.ctypes.data, s, t, tlut, tlut_type, lut_offset, addr_xor, out.ctypes.data)
    return tuple(int(v) for v in out)


def pin_texel_formats(L):
    r = Rdp(L)
    try:
        W, Hh = 32, 64  # 2048 16-bit texels per TMEM fill, odd and even lines
        for fmt, conv in ((TF_RGBA, _rgba16), (TF_IA, _ia16)):
            for base in range(0, 65536, W * Hh):
                tile = _tile(fmt, 2, stride=2 * W)
                for t in range(Hh):
                    for s in range(W):
                        _put16(r.tmem, _texel_byte(0, 2 * W, s, t, 16), base + t * W + s)
                for t in range(Hh):
                    for s in range(W):
                        v = base + t * W + s
                        assert _fetch(L, r, tile, s, t) == conv(v), (fmt, s, t, hex(v))
        # 8-bit and 4-bit formats over every value, on odd and even lines.
        ex4 = lambda c: c * 17  # noqa: E731
        ex3 = lambda c: (c << 5) | (c << 2) | (c >> 1)  # noqa: E731
        eight = {
            (TF_IA, 1): lambda v: (ex4(v >> 4),) * 3 + (ex4(v & 15),),
            (TF_I, 1): lambda v: (v,) * 4,
            (TF_RGBA, 1): lambda v: (v,) * 4,
            (TF_CI, 1): lambda v: (v,) * 4,
        }
        for (fmt, size), conv in eight.items():
            tile = _tile(fmt, size, stride=32)
            for t in range(16):
                for s in range(32):
                    _put8(r.tmem, _texel_byte(0, 32, s, t, 8), (t * 32 + s) & 0xFF)
            for t in range(16):
                for s in range(32):
                    assert _fetch(L, r, tile, s, t) == conv((t * 32 + s) & 0xFF), (fmt, size, s, t)
        four = {
            TF_IA: lambda v, p: (ex3(v >> 1),) * 3 + (255 if v & 1 else 0,),
            TF_I: lambda v, p: (ex4(v),) * 4,
            TF_RGBA: lambda v, p: (ex4(v),) * 4,
            TF_CI: lambda v, p: ((p << 4) | v,) * 4,
        }
        for fmt, conv in four.items():
            for pal in (0, 5, 15):
                tile = _tile(fmt, 0, stride=16, palette=pal)
                for t in range(8):
                    for s in range(32):
                        b = _texel_byte(0, 16, s, t, 4)
                        v = (s + 3 * t) & 15
                        old = r.tmem[b ^ 3]
                        r.tmem[b ^ 3] = (old & 0x0F) | (v << 4) if s % 2 == 0 else (old & 0xF0) | v
                for t in range(8):
                    for s in range(32):
                        assert _fetch(L, r, tile, s, t) == conv((s + 3 * t) & 15, pal), (fmt, pal, s, t)
        # RGBA32: red/green in the low half, blue/alpha at the same place in the high half.
        rng = np.random.default_rng(5)
        tile = _tile(TF_RGBA, 3, stride=64)
        for _ in range(4):
            vals = rng.integers(0, 1 << 32, size=(32, 32), dtype=np.uint64)
            for t in range(32):
                for s in range(32):
                    b = _texel_byte(0, 64, s, t, 16, 0x7FF)
                    v = int(vals[t, s])
                    _put16(r.tmem, b, v >> 16)
                    _put16(r.tmem, b + 0x800, v & 0xFFFF)
            for t in range(32):
                for s in range(32):
                    v = int(vals[t, s])
                    assert _fetch(L, r, tile, s, t) == (v >> 24, (v >> 16) & 255, (v >> 8) & 255, v & 255)
        # TLUT: colour indices look up 16-bit entries in the high half, each
        # entry stored four times (one copy per bank).
        for tlut_type, conv in ((0, _rgba16), (1, _ia16)):
            lut = [int(v) for v in rng.integers(0, 65536, 256)]
            for i, e in enumerate(lut):
                for bank in range(4):
                    _put16(r.tmem, 0x800 + 8 * i + 2 * bank, e)
            tile8 = _tile(TF_CI, 1, stride=32)
            for t in range(8):
                for s in range(32):
                    _put8(r.tmem, _texel_byte(0, 32, s, t, 8, 0x7FF), (t * 32 + s) & 255)
            for t in range(8):
                for s in range(32):
                    want = conv(lut[(t * 32 + s) & 255])
                    for off, xor in ((0, 1), (1, 1), (2, 2), (3, 2)):
                        assert _fetch(L, r, tile8, s, t, 1, tlut_type, off, xor) == want
            for pal in (0, 9):
                tile4 = _tile(TF_CI, 0, stride=16, palette=pal)
                for t in range(4):
                    for s in range(32):
                        b = _texel_byte(0, 16, s, t, 4, 0x7FF)
                        v = (s * 7 + t) & 15
                        old = r.tmem[b ^ 3]
                        r.tmem[b ^ 3] = (old & 0x0F) | (v << 4) if s % 2 == 0 else (old & 0xF0) | v
                for t in range(4):
                    for s in range(32):
                        assert _fetch(L, r, tile4, s, t, 1, tlut_type, 0, 1) == conv(lut[(pal << 4) | ((s * 7 + t) & 15)])
        # The four copies of an entry are four banks, one per filter tap: tap k
        # reads bank k for the lower triangle of the texel square (address
        # xor 1) and bank k ^ 3 for the upper one (xor 2).
        for i in range(256):
            for bank in range(4):
                _put16(r.tmem, 0x800 + 8 * i + 2 * bank, (i << 8) | (bank << 4) | 1)
        tile8 = _tile(TF_CI, 1, stride=32)
        _put8(r.tmem, _texel_byte(0, 32, 3, 0, 8, 0x7FF), 77)
        for off in range(4):
            for xor, bank in ((1, off), (2, off ^ 3)):
                got = _fetch(L, r, tile8, 3, 0, 1, 1, off, xor)
                assert got == _ia16((77 << 8) | (bank << 4) | 1), (off, xor, got)
        # YUV: chroma pairs (U, V) in the low half, one per two texels; luma in the high half.
        tile = _tile(TF_YUV, 2, stride=32)
        y_img = rng.integers(0, 256, size=(8, 32))
        uv_img = rng.integers(0, 65536, size=(8, 16))
        for t in range(8):
            for s in range(32):
                _put8(r.tmem, 0x800 + _texel_byte(0, 32, s, t, 8, 0x7FF), int(y_img[t, s]))
            for cx in range(16):
                _put16(r.tmem, _texel_byte(0, 32, cx, t, 16, 0x7FF), int(uv_img[t, cx]))
        for t in range(8):
            for s in range(32):
                out = np.zeros(4, np.int32)
                L.rdpo_test_fetch_yuv(r.h, tile.ctypes.data, s, t, s >> 1, out.ctypes.data)
                uv = int(uv_img[t, s >> 1])
                assert tuple(out) == ((uv >> 8) - 128, (uv & 255) - 128, int(y_img[t, s]), int(y_img[t, s]))
    finally:
        r.close()


# ---------------------------------------------------------------------------
# Addressing: shift, clamp, wrap and mirror as arithmetic.
# ---------------------------------------------------------------------------
def _model_texel_coord(c, lo, hi, shift, clamp, mask, mirror):
    """Texel index of coordinate c (1/32 texel) on a tile [lo, hi] (quarter
    texels): shift right by `shift` (or left by 16 - shift from 11 on, in 16
    bits), clamp to the tile's texels, then reduce modulo 2^mask, reflecting
    every odd period when mirrored."""
    c = min(max(c, -0x8000), 0x7FFF)
    c = c >> shift if shift < 11 else _sext(c << (16 - shift), 16)
    if clamp:
        if c >= 8 * hi:
            c = (((hi >> 2) - (lo >> 2)) & 0x3FF) << 5
        else:
            c = max(c - 8 * lo, 0)
    else:
        c -= 8 * lo
    s = c >> 5
    if mask:
        q, rem = divmod(s, 1 << mask)
        s = (1 << mask) - 1 - rem if (mirror and q & 1) else rem
    return s


def pin_texture_addressing(L):
    r = Rdp(L)
    try:
        n = 16
        img = (np.arange(n * n).reshape(n, n) * 37 + 11) & 0xFF  # distinct I8 texels, 16 x 16
        for t in range(n):
            for s in range(n):
                _put8(r.tmem, _texel_byte(0, n, s, t, 8), int(img[t, s]))
        zero4 = _arr([0, 0, 0, 0])
        rng = np.random.default_rng(9)
        for case in range(60):
            mask = int(rng.integers(0, 5))
            mirror = bool(rng.random() < 0.5) and mask > 0
            clamp = bool(rng.random() < 0.5) or mask == 0
            shift = int(rng.choice([0, 1, 2, 11, 14, 15]))
            lo = int(rng.integers(0, 6)) * 4 + int(rng.integers(0, 4))
            hi = lo + int(rng.integers(4, 4 * (n - lo // 4 - 1)))
            flags = (TILE_CLAMP_S if clamp else 0) | (TILE_MIRROR_S if mirror else 0)
            tile = _tile(TF_I, 1, stride=n, mask_s=mask, shift_s=shift, slo=lo, shi=hi, mask_t=mask, shift_t=shift,
                         tlo=lo, thi=hi, flags=flags | (flags << 2))
            for c in range(-3000, 3000, 13):
                st = _arr([c, c])
                out = np.zeros(4, np.int32)
                L.rdpo_test_sample_texture(r.h, tile.ctypes.data, st.ctypes.data, 0, 0, 0, 0, 0, 1, zero4.ctypes.data,
                                           zero4.ctypes.data, out.ctypes.data)
                s = _model_texel_coord(c, lo, hi, shift, clamp, mask, mirror)
                tt = _model_texel_coord(c, lo, hi, shift, clamp, mask, mirror)
                if not (0 <= s < n and 0 <= tt < n):
                    continue
                assert out[0] == img[tt, s], (case, c, mask, mirror, clamp, shift, lo, hi, s, tt, out[0])
    finally:
        r.close()


# ---------------------------------------------------------------------------
# Filtering: constant textures and linear ramps.
# ---------------------------------------------------------------------------
def _sample(L, r, tile, st, quad=1, mid=0, bilerp=1, tlut=0, tlut_type=0):
    out = np.zeros(4, np.int32)
    z = _arr([0, 0, 0, 0])
    L.rdpo_test_sample_texture(r.h, tile.ctypes.data, _arr(st).ctypes.data, tlut, tlut_type, quad, mid, 0, bilerp,
                               z.ctypes.data, z.ctypes.data, out.ctypes.data)
    return out


def pin_filtering(L):
    r = Rdp(L)
    try:
        rng = np.random.default_rng(17)
        n = 32
        tile = _tile(TF_I, 1, stride=n)
        for K in (0, 1, 128, 254, 255):
            for t in range(n):
                for s in range(n):
                    _put8(r.tmem, _texel_byte(0, n, s, t, 8), K)
            for _ in range(200):
                st = [int(v) for v in rng.integers(0, 30 * 32, 2)]
                assert list(_sample(L, r, tile, st)) == [K] * 4
                assert list(_sample(L, r, tile, [st[0] | 16, st[1] | 16], mid=1)) == [K] * 4
        for _ in range(12):
            b, c = int(rng.integers(-3, 4)), int(rng.integers(-3, 4))
            a = 128 - 15 * (b + c)  # every texel of the 32 x 32 ramp stays in 35..221
            f = lambda s, t: a + b * s + c * t  # noqa: E731
            for t in range(n):
                for s in range(n):
                    _put8(r.tmem, _texel_byte(0, n, s, t, 8), f(s, t) & 0xFF)
            for _ in range(300):
                s0, t0 = int(rng.integers(0, n - 1)), int(rng.integers(0, n - 1))
                fx, fy = int(rng.integers(0, 32)), int(rng.integers(0, 32))
                got = _sample(L, r, tile, [s0 * 32 + fx, t0 * 32 + fy])
                want = math.floor(f(s0 + Fraction(fx, 32), t0 + Fraction(fy, 32)) + Fraction(1, 2))
                assert list(got) == [want] * 4, (a, b, c, s0, t0, fx, fy, list(got), want)
                got = _sample(L, r, tile, [s0 * 32 + 16, t0 * 32 + 16], mid=1)
                want = math.floor(f(s0 + Fraction(1, 2), t0 + Fraction(1, 2)) + Fraction(1, 2))
                assert list(got) == [want] * 4
        # Through a TLUT of IA16 entries (intensity = alpha = index): a ramp of indices.
        for i in range(256):
            for bank in range(4):
                _put16(r.tmem, 0x800 + 8 * i + 2 * bank, (i << 8) | i)
        tile = _tile(TF_CI, 1, stride=n)
        for t in range(n):
            for s in range(n):
                _put8(r.tmem, _texel_byte(0, n, s, t, 8, 0x7FF), 40 + 2 * s + 3 * t)
        for _ in range(300):
            s0, t0 = int(rng.integers(0, n - 1)), int(rng.integers(0, n - 1))
            fx, fy = int(rng.integers(0, 32)), int(rng.integers(0, 32))
            got = _sample(L, r, tile, [s0 * 32 + fx, t0 * 32 + fy], tlut=1, tlut_type=1)
            want = math.floor(40 + 2 * (s0 + Fraction(fx, 32)) + 3 * (t0 + Fraction(fy, 32)) + Fraction(1, 2))
            assert list(got) == [want] * 4
    finally:
        r.close()


# ---------------------------------------------------------------------------
# Level of detail.
# ---------------------------------------------------------------------------
def pin_lod(L):
    out = np.zeros(3, np.int32)
    for max_level in (1, 3, 5):
        for base in (0, 2, 6):
            for d in list(range(1, 32)) + list(range(32, 0x4000, 7)) + [32 << k for k in range(9)]:
                stq = _arr([0, 0, d, 0, 0, d // 2])
                L.rdpo_test_compute_lod(base, max_level, 0, stq.ctypes.data, 0, 1, 0, 0, out.ctypes.data)
                if d < 32:  # magnified: the base level, no fraction
                    want = (base, base, 0)
                else:
                    k = (d // 32).bit_length() - 1  # floor(log2(texels per pixel))
                    frac = math.floor(256 * (Fraction(d, 32 << k) - 1))
                    if k >= max_level:  # beyond the last level
                        want = ((base + max_level) & 7, (base + max_level) & 7, 0xFF)
                    else:
                        want = ((base + k) & 7, (base + k + 1) & 7, frac)
                assert tuple(out) == want, (max_level, base, d, tuple(out), want)


# ---------------------------------------------------------------------------
# Combiner input selection (the RDP's colour combiner tables).
# ---------------------------------------------------------------------------
def pin_combiner_selection(L):
    rng = np.random.default_rng(23)
    for _ in range(20):
        vals = [int(v) for v in rng.integers(0, 0x1FF, 18)]
        k = [int(v) for v in rng.integers(0, 0xFF, 16)]
        shade, comb, t0, t1 = vals[0:4], vals[4:8], vals[8:12], vals[12:16]
        lod, noise = vals[16], vals[17]
        named = {"combined": comb, "texel0": t0, "texel1": t1, "shade": shade}
        rgb = {
            0: {0: "combined", 1: "texel0", 2: "texel1", 4: "shade", 6: ("one",), 7: ("noise",)},
            1: {0: "combined", 1: "texel0", 2: "texel1", 4: "shade"},
            2: {0: "combined", 1: "texel0", 2: "texel1", 4: "shade", 7: ("alpha", "combined"),
                8: ("alpha", "texel0"), 9: ("alpha", "texel1"), 11: ("alpha", "shade"), 13: ("lod",)},
            3: {0: "combined", 1: "texel0", 2: "texel1", 4: "shade", 6: ("one",)},
        }
        alpha = {
            2: {0: ("lod",), 1: "texel0", 2: "texel1", 4: "shade"},
            0: {0: "combined", 1: "texel0", 2: "texel1", 4: "shade", 6: ("one",)},
        }
        alpha[1] = alpha[3] = alpha[0]
        special = {(1, 7), (2, 15)}  # K4 and K5 use the constant slots' packed form
        for which in range(4):
            kslot = k[4 * which: 4 * which + 4]
            for sel in range(32 if which == 2 else 16):
                if (which, sel) in special:
                    continue
                for sel_a in range(8):
                    out = np.zeros(4, np.int32)
                    L.rdpo_test_select(_arr(vals).ctypes.data, _arr(k).ctypes.data, which, sel, sel_a,
                                       out.ctypes.data)
                    src = rgb[which].get(sel)
                    if src is None:
                        want = kslot[0:3]
                    elif isinstance(src, str):
                        want = named[src][0:3]
                    elif src[0] == "one":
                        want = [0x100] * 3
                    elif src[0] == "noise":
                        want = [noise] * 3
                    elif src[0] == "lod":
                        want = [lod] * 3
                    else:
                        want = [named[src[1]][3]] * 3
                    a = alpha[which].get(sel_a)
                    if a is None:
                        want_a = kslot[3]
                    elif isinstance(a, str):
                        want_a = named[a][3]
                    elif a[0] == "one":
                        want_a = 0x100
                    else:
                        want_a = lod
                    assert list(out) == list(want) + [want_a], (which, sel, sel_a, list(out), want, want_a)


# ---------------------------------------------------------------------------
# Depth test invariants.
# ---------------------------------------------------------------------------
def _dt(L, z, dz, dzc, depth, mdz, cvg, mcvg, mode, aa=1, compare=1, force=0):
    out = np.zeros(5, np.int32)
    p = L.rdpo_test_depth_test(z, dz, dzc, depth, mdz, cvg, mcvg, compare, mode, force, aa, out.ctypes.data)
    return p, out


def pin_depth_test(L):
    rng = np.random.default_rng(31)
    for _ in range(300):
        depth = int(rng.integers(0, 0x3FFF))
        mdz = int(rng.integers(0, 16))
        dzc = int(rng.integers(0, 16))
        dz = 1 << dzc
        mz = L.rdpo_test_z_decompress(depth)
        cvg, mcvg = int(rng.integers(1, 9)), int(rng.integers(0, 8))
        zs = range(max(mz - 3000, 0), min(mz + 3000, 0x3FFFF), 3)
        # no compare: everything passes
        assert all(_dt(L, z, dz, dzc, depth, mdz, cvg, mcvg, 0, compare=0)[0] for z in zs[::50])
        # opaque: monotone in z, and anything in front passes
        res = [_dt(L, z, dz, dzc, depth, mdz, cvg, mcvg, 0)[0] for z in zs]
        for i in range(1, len(res)):
            assert res[i] <= res[i - 1], ("opaque not monotone", depth, mdz, dz)
        assert all(p for z, p in zip(zs, res) if z < mz)
        # transparent: exactly the fragments in front
        for z in zs[::7]:
            assert _dt(L, z, dz, dzc, depth, mdz, cvg, mcvg, 2)[0] == (z < mz)
        # decal: a window around the stored depth, the same width on both sides
        def decal(z):
            return _dt(L, z, dz, dzc, depth, mdz, cvg, mcvg, 3)[0]

        if mz < 0x3FFFF:
            assert decal(mz)
            lo, hi = 0, mz  # first passing z below mz
            while lo < hi:
                mid = (lo + hi) // 2
                lo, hi = (lo, mid) if decal(mid) else (mid + 1, hi)
            first = lo
            lo, hi = mz, 0x3FFFF  # last passing z above mz
            while lo < hi:
                mid = (lo + hi + 1) // 2
                lo, hi = (mid, hi) if decal(mid) else (lo, mid - 1)
            last = lo
            for z in (first, (first + mz) // 2, (mz + last) // 2, last):
                assert decal(z)
            if first > 0:
                assert not decal(first - 1)
            if last < 0x3FFFE:
                assert not decal(last + 1)
                if first > 0:
                    assert mz - first == last - mz, ("decal window not symmetric", mz, first, last)
        # interpenetrating: the opaque result wherever the overlap case does not apply
        for z in zs[::11]:
            p1, o1 = _dt(L, z, dz, dzc, depth, mdz, cvg, mcvg, 1)
            p0, o0 = _dt(L, z, dz, dzc, depth, mdz, cvg, mcvg, 0)
            if not (z < mz and cvg + mcvg >= 8):
                assert p1 == p0
            else:
                assert p1 and 0 <= o1[0] <= 8
            assert 0 <= o0[3] <= 4 and 0 <= o0[4] <= 4
    # the largest stored depth passes in every mode but decal
    for mode, want in ((0, 1), (1, 1), (2, 1), (3, 0)):
        depth = 0x3FFF  # decompresses to 0x3ffff
        assert L.rdpo_test_z_decompress(depth) == 0x3FFFF
        assert _dt(L, 0x3FFFF, 1, 0, depth, 0, 4, 3, mode)[0] == want


# ---------------------------------------------------------------------------
# Blender.
# ---------------------------------------------------------------------------
def _blend(L, r, pixel, memory, modes, final, blend_en=1, force=0, color_on_cvg=0, wrap=0, fog=(0, 0, 0, 0),
           blendc=(0, 0, 0, 0), shade_alpha=0):
    out = np.zeros(3, np.int32)
    inp = _arr(list(pixel) + list(memory) + list(fog) + list(blendc) + [shade_alpha])
    L.rdpo_test_blender(r.h, inp.ctypes.data, _arr(modes).ctypes.data, force, blend_en, color_on_cvg, wrap, 0, 0,
                        final, out.ctypes.data)
    return list(out)


def pin_blender(L):
    r = Rdp(L)
    try:
        rng = np.random.default_rng(41)
        modes = (0, 0, 1, 0)  # pixel * alpha + memory * (1 - alpha)
        for _ in range(3000):
            p = [int(v) for v in rng.integers(0, 256, 3)] + [int(rng.integers(0, 256))]
            m = [int(v) for v in rng.integers(0, 256, 3)] + [int(rng.integers(0, 256))]
            w = p[3] >> 3  # 5-bit weight of the pixel; the memory weight is 32 - w
            want = [(p[i] * w + m[i] * (32 - w)) >> 5 for i in range(3)]
            assert _blend(L, r, p, m, modes, final=0) == want, (p, m, want)
            if p[3] != 255:
                assert _blend(L, r, p, m, modes, final=1) == want, (p, m)
            assert _blend(L, r, p, m, modes, final=1, blend_en=0) == p[:3]
            assert _blend(L, r, p, m, modes, final=1, color_on_cvg=1) == m[:3]
        for _ in range(200):
            p = [int(v) for v in rng.integers(0, 256, 3)] + [255]
            m = [int(v) for v in rng.integers(0, 256, 4)]
            assert _blend(L, r, p, m, modes, final=1) == p[:3]  # an opaque pixel replaces memory
    finally:
        r.close()


def pin_blend_coverage(L):
    for cvg in range(9):
        for mem in range(8):
            for en in (0, 1):
                assert L.rdpo_test_blend_coverage(cvg, mem, en, 0) == (min(7, mem + cvg) if en else (cvg - 1) % 8)
                assert L.rdpo_test_blend_coverage(cvg, mem, en, 1) == (cvg + mem) % 8
                assert L.rdpo_test_blend_coverage(cvg, mem, en, 2) == 7
                assert L.rdpo_test_blend_coverage(cvg, mem, en, 3) == mem


# ---------------------------------------------------------------------------
# Framebuffer formats, fill and copy.
# ---------------------------------------------------------------------------
def _rd16(r, i):
    return (int(r.rdram[(2 * i) ^ 3]) << 8) | int(r.rdram[(2 * i + 1) ^ 3])


def _wr16(r, i, v):
    r.rdram[(2 * i) ^ 3] = (v >> 8) & 0xFF
    r.rdram[(2 * i + 1) ^ 3] = v & 0xFF


def _load(L, r, fmt, i):
    out = np.zeros(4, np.int32)
    L.rdpo_test_load_color(r.h, fmt, i, out.ctypes.data)
    return [int(v) for v in out]


def _store(L, r, fmt, i, c):
    L.rdpo_test_store_color(r.h, fmt, i, _arr(c).ctypes.data)


def pin_framebuffer_formats(L):
    r = Rdp(L)
    try:
        # RGBA5551: 5-bit channels in the top bits; the 3-bit coverage is the
        # alpha bit above the two hidden bits.  Every value, every hidden value.
        for w in range(65536):
            i = 1000 + (w & 255)
            h = w >> 14
            _wr16(r, i, w)
            r.hidden[i] = h
            c = _load(L, r, FB_RGBA5551, i)
            assert c == [((w >> 11) & 31) << 3, ((w >> 6) & 31) << 3, ((w >> 1) & 31) << 3,
                         ((w & 1) << 7) | (h << 5)], (hex(w), h, c)
            _wr16(r, i, 0)
            r.hidden[i] = 0
            _store(L, r, FB_RGBA5551, i, c)
            assert _rd16(r, i) == w and r.hidden[i] == h, (hex(w), h)
        # IA88: intensity byte, alpha byte; the hidden bits repeat alpha's low bit.
        for w in range(0, 65536, 3):
            _wr16(r, 2000, w)
            c = _load(L, r, FB_IA88, 2000)
            assert c == [w >> 8, w >> 8, w >> 8, w & 255]
            _wr16(r, 2000, 0)
            _store(L, r, FB_IA88, 2000, c)
            assert _rd16(r, 2000) == w and r.hidden[2000] == 3 * (w & 1)
        # RGBA8888: one word, the hidden bits of its halfwords repeat G's and A's low bits.
        rng = np.random.default_rng(3)
        for v in rng.integers(0, 1 << 32, size=4000, dtype=np.uint64):
            v = int(v)
            words = r.rdram.view(np.uint32)
            words[700] = v
            c = _load(L, r, FB_RGBA8888, 700)
            assert c == [v >> 24, (v >> 16) & 255, (v >> 8) & 255, v & 255]
            words[700] = 0
            _store(L, r, FB_RGBA8888, 700, c)
            assert int(words[700]) == v
            assert r.hidden[1400] == 3 * ((v >> 16) & 1) and r.hidden[1401] == 3 * (v & 1)
        # I8: one byte per pixel; an odd pixel's byte writes the pair's hidden bits.
        for i in (5000, 5001):
            for v in range(256):
                r.rdram[i ^ 3] = v
                c = _load(L, r, FB_I8, i)
                assert c[:3] == [v] * 3
                r.rdram[i ^ 3] = 0
                _store(L, r, FB_I8, i, [v, v, v, v])
                assert r.rdram[i ^ 3] == v
                if i & 1:
                    assert r.hidden[i >> 1] == 3 * (v & 1)
        # Depth: 14-bit compressed depth above 2 bits of dz; 2 more dz bits hidden.
        out = np.zeros(2, np.uint32)
        for w in range(0, 65536, 7):
            for h in range(4):
                _wr16(r, 3000, w)
                r.hidden[3000] = h
                L.rdpo_test_load_depth(r.h, 3000, out.ctypes.data)
                assert (int(out[0]), int(out[1])) == (w >> 2, ((w & 3) << 2) | h)
                _wr16(r, 3000, 0)
                r.hidden[3000] = 0
                L.rdpo_test_store_depth(r.h, 3000, int(out[0]), int(out[1]))
                assert _rd16(r, 3000) == w and r.hidden[3000] == h
    finally:
        r.close()


def pin_fill_and_copy(L):
    r = Rdp(L)
    try:
        rng = np.random.default_rng(61)
        io = np.zeros(6, np.int32)
        # Fill: the 32-bit fill colour tiles memory; a 16-bit pixel at an even
        # index takes its upper half, at an odd index its lower half.
        for col in [int(v) for v in rng.integers(0, 1 << 32, 200, dtype=np.uint64)]:
            for i in range(4000, 4008):
                io[:] = 0
                L.rdpo_test_fill_color(FB_RGBA5551, i, 0, col, io.ctypes.data)
                _store(L, r, FB_RGBA5551, i, io[:4])
                half = (col >> 16) if i % 2 == 0 else (col & 0xFFFF)
                assert _rd16(r, i) == half and r.hidden[i] == 3 * (half & 1), (hex(col), i)
                io[:] = 0
                L.rdpo_test_fill_color(FB_IA88, i, 0, col, io.ctypes.data)
                _store(L, r, FB_IA88, i, io[:4])
                assert _rd16(r, i) == half
            for i in range(8000, 8008):
                io[:] = 0
                L.rdpo_test_fill_color(FB_I8, i, 0, col, io.ctypes.data)
                _store(L, r, FB_I8, i, io[:4])
                assert r.rdram[i ^ 3] == (col >> (8 * (3 - (i & 3)))) & 0xFF
            io[:] = 0
            L.rdpo_test_fill_color(FB_RGBA8888, 900, 0, col, io.ctypes.data)
            _store(L, r, FB_RGBA8888, 900, io[:4])
            assert int(r.rdram.view(np.uint32)[900]) == col
        # Copy: a 16-bit texel word lands in memory unchanged.
        for word in range(65536):
            io[:] = 0
            L.rdpo_test_copy_pipeline(FB_RGBA5551, 0, word, io.ctypes.data)
            _store(L, r, FB_RGBA5551, 6000, io[:4])
            assert _rd16(r, 6000) == word and r.hidden[6000] == 3 * (word & 1), hex(word)
        for word in range(256):
            io[:] = 0
            L.rdpo_test_copy_pipeline(FB_I8, 0, word | 0x5500, io.ctypes.data)
            _store(L, r, FB_I8, 7002, io[:4])
            assert r.rdram[7002 ^ 3] == word
    finally:
        r.close()


PINS = {
    "span_walker": pin_span_walker,
    "interpolation_direct_form": pin_interpolation_direct_form,
    "interpolation_plane": pin_interpolation_plane,
    "texel_formats": pin_texel_formats,
    "texture_addressing": pin_texture_addressing,
    "filtering": pin_filtering,
    "lod": pin_lod,
    "combiner_selection": pin_combiner_selection,
    "depth_test": pin_depth_test,
    "blender": pin_blender,
    "blend_coverage": pin_blend_coverage,
    "framebuffer_formats": pin_framebuffer_formats,
    "fill_and_copy": pin_fill_and_copy,
}

# Planted mistake (RDPO_PLANT value in rdp_oracle.c) -> the pin that must catch it.
PLANTS = {
    1: "span_walker",                # quantised edges lose their sticky bit
    2: "span_walker",                # H/M edges start at YH instead of its whole scanline
    3: "interpolation_direct_form",  # x fraction taken one bit too high
    4: "interpolation_plane",        # centroid x offset ignores the sub-line
    5: "interpolation_plane",        # z keeps one fraction bit too many
    6: "texel_formats",              # RGBA16 channels not bit-replicated
    7: "texel_formats",              # IA4 intensity expansion drops a term
    8: "texel_formats",              # odd TMEM lines not word-swapped
    9: "texture_addressing",         # mirror reflects one texel off
    10: "texture_addressing",        # clamp to the tile's edge one step late
    11: "filtering",                 # bilinear rounding constant off by one
    12: "lod",                       # level chosen one step too fine
    13: "combiner_selection",        # shade alpha selector reads texel 0 alpha
    14: "depth_test",                # nearer test strict
    15: "blender",                   # memory weight one short
    16: "framebuffer_formats",       # RGBA5551 alpha bit from the wrong coverage bit
    17: "fill_and_copy",             # fill halves swapped
    18: "fill_and_copy",             # copy alpha from the wrong bit
    19: "blend_coverage",            # wrap mode wraps at 16
    20: "texel_formats",             # TLUT bank ignores the address xor
}


@pytest.fixture(scope="module")
def lib(oracle_mod):
    return _bind(oracle_mod._load())


@pytest.mark.parametrize("name", sorted(PINS))
def test_pin(lib, name):
    PINS[name](lib)


def _planted(n):
    out_dir = os.path.join(ORACLE, "build", "planted")
    os.makedirs(out_dir, exist_ok=True)
    out = os.path.join(out_dir, f"librdp_oracle_plant{n}.so")
    srcs = [os.path.join(ORACLE, f) for f in ("rdp_oracle.c", "vi_oracle.c", "rdp_oracle.h", "vi_oracle.h")]
    if not os.path.exists(out) or os.path.getmtime(out) < max(os.path.getmtime(s) for s in srcs):
        subprocess.run(["gcc", "-O1", "-fwrapv", "-fPIC", "-std=c99", "-shared", f"-DRDPO_PLANT={n}", "-o", out,
                        srcs[0], srcs[1], "-lm"], check=True)
    return _bind(ctypes.CDLL(out))


@pytest.mark.parametrize("plant", sorted(PLANTS))
def test_planted_mistake_fails_its_pin(plant):
    L = _planted(plant)
    with pytest.raises(AssertionError):
        PINS[PLANTS[plant]](L)
