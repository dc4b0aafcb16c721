"""Seeded RDP command-stream builder for the hot-path workloads.

Plays the role of the reference's CommandBuilder and triangle converter
(rdp_command_builder.cpp:278-419, triangle_converter.cpp:129-273): it turns
screen-space vertices into raw RDP command words.  Triangle setup is
vectorised over whole batches with numpy; edge setup is integer-exact (quarter
pixel vertices, round-away-from-zero slopes masked with ~7, sub-scanline
start fixup) and attribute gradients follow the same plane equations,
quantised to the RDP's 16.16 formats.

Everything here produces input data only.  Rendering semantics live in the
library and in the oracle.
"""
from __future__ import annotations

import dataclasses
import math
from typing import Dict, List, Optional, Sequence

import numpy as np

# ----------------------------------------------------------------------------
# Opcodes (rdp_common.hpp:37-81)
# ----------------------------------------------------------------------------
OP_FILL_TRIANGLE = 0x08
OP_SHADE_TRIANGLE = 0x0C
OP_SHADE_Z_TRIANGLE = 0x0D
OP_SHADE_TEX_TRIANGLE = 0x0E
OP_SHADE_TEX_Z_TRIANGLE = 0x0F
OP_TEX_Z_TRIANGLE = 0x0B
OP_TEXTURE_RECTANGLE = 0x24
OP_TEXTURE_RECTANGLE_FLIP = 0x25
OP_SYNC_LOAD = 0x26
OP_SYNC_PIPE = 0x27
OP_SYNC_TILE = 0x28
OP_SYNC_FULL = 0x29
OP_SET_CONVERT = 0x2C
OP_SET_SCISSOR = 0x2D
OP_SET_PRIM_DEPTH = 0x2E
OP_SET_OTHER_MODES = 0x2F
OP_LOAD_TLUT = 0x30
OP_SET_TILE_SIZE = 0x32
OP_LOAD_BLOCK = 0x33
OP_LOAD_TILE = 0x34
OP_SET_TILE = 0x35
OP_FILL_RECTANGLE = 0x36
OP_SET_FILL_COLOR = 0x37
OP_SET_FOG_COLOR = 0x38
OP_SET_BLEND_COLOR = 0x39
OP_SET_PRIM_COLOR = 0x3A
OP_SET_ENV_COLOR = 0x3B
OP_SET_COMBINE = 0x3C
OP_SET_TEXTURE_IMAGE = 0x3D
OP_SET_MASK_IMAGE = 0x3E
OP_SET_COLOR_IMAGE = 0x3F

# Texture formats / sizes (rdp_common.hpp:176-190)
FMT_RGBA, FMT_YUV, FMT_CI, FMT_IA, FMT_I = 0, 1, 2, 3, 4
SIZ_4, SIZ_8, SIZ_16, SIZ_32 = 0, 1, 2, 3

# VI registers (rdp_common.hpp:324-341)
VI_CONTROL, VI_ORIGIN, VI_WIDTH, VI_INTR, VI_V_CURRENT_LINE, VI_TIMING, VI_V_SYNC, \
    VI_H_SYNC, VI_LEAP, VI_H_START, VI_V_START, VI_V_BURST, VI_X_SCALE, VI_Y_SCALE = range(14)

RDRAM_SIZE = 8 * 1024 * 1024

COMMAND_LENGTHS = np.zeros(64, dtype=np.int32)
COMMAND_LENGTHS[0] = 2
COMMAND_LENGTHS[0x08:0x10] = [8, 12, 24, 28, 24, 28, 40, 44]
COMMAND_LENGTHS[0x24:0x40] = 2
COMMAND_LENGTHS[0x24] = 4
COMMAND_LENGTHS[0x25] = 4
COMMAND_LENGTHS[0x31] = 0


def _u32(v) -> int:
    return int(v) & 0xFFFFFFFF


@dataclasses.dataclass
class OtherModes:
    """Fields of SetOtherModes (rdp_command_builder.cpp:143-205 layout)."""
    cycle_type: int = 0          # 0 1-cycle, 1 2-cycle, 2 copy, 3 fill
    perspective: bool = False
    detail_lod: bool = False
    sharpen_lod: bool = False
    tex_lod: bool = False
    tlut: bool = False
    tlut_ia: bool = False
    sample_quad: bool = False
    mid_texel: bool = False
    bilerp0: bool = False
    bilerp1: bool = False
    convert_one: bool = False
    rgb_dither: int = 3          # 0 magic, 1 bayer, 2 noise, 3 off
    alpha_dither: int = 3        # 0 pattern, 1 inv pattern, 2 noise, 3 off
    blend: Sequence[Sequence[int]] = ((0, 0, 0, 0), (0, 0, 0, 0))   # per cycle: 1a, 1b, 2a, 2b
    force_blend: bool = False
    alpha_cvg_select: bool = False
    cvg_times_alpha: bool = False
    z_mode: int = 0
    coverage_mode: int = 0
    color_on_coverage: bool = False
    image_read: bool = False
    depth_write: bool = False
    depth_test: bool = False
    aa: bool = False
    prim_depth: bool = False
    alpha_test_dither: bool = False
    alpha_test: bool = False

    def words(self) -> List[int]:
        w0 = OP_SET_OTHER_MODES << 24
        w0 |= (self.cycle_type & 3) << 20
        for bit, flag in ((19, self.perspective), (18, self.detail_lod), (17, self.sharpen_lod),
                          (16, self.tex_lod), (15, self.tlut), (14, self.tlut_ia), (13, self.sample_quad),
                          (12, self.mid_texel), (11, self.bilerp0), (10, self.bilerp1), (9, self.convert_one)):
            if flag:
                w0 |= 1 << bit
        w0 |= (self.rgb_dither & 3) << 6
        w0 |= (self.alpha_dither & 3) << 4
        b0, b1 = self.blend
        w1 = ((b0[0] & 3) << 30) | ((b1[0] & 3) << 28) | ((b0[1] & 3) << 26) | ((b1[1] & 3) << 24)
        w1 |= ((b0[2] & 3) << 22) | ((b1[2] & 3) << 20) | ((b0[3] & 3) << 18) | ((b1[3] & 3) << 16)
        for bit, flag in ((14, self.force_blend), (13, self.alpha_cvg_select), (12, self.cvg_times_alpha),
                          (7, self.color_on_coverage), (6, self.image_read), (5, self.depth_write),
                          (4, self.depth_test), (3, self.aa), (2, self.prim_depth),
                          (1, self.alpha_test_dither), (0, self.alpha_test)):
            if flag:
                w1 |= 1 << bit
        w1 |= (self.z_mode & 3) << 10
        w1 |= (self.coverage_mode & 3) << 8
        return [_u32(w0), _u32(w1)]


def combine_words(c0_rgb, c0_alpha, c1_rgb, c1_alpha) -> List[int]:
    """SetCombine from (muladd, mulsub, mul, add) selectors per cycle.

    Bit layout inverts rdp_device.cpp:661-684."""
    a0, b0, c0, d0 = c0_rgb
    aa0, ab0, ac0, ad0 = c0_alpha
    a1, b1, c1, d1 = c1_rgb
    aa1, ab1, ac1, ad1 = c1_alpha
    w0 = OP_SET_COMBINE << 24
    w0 |= (a0 & 0xF) << 20 | (c0 & 0x1F) << 15 | (aa0 & 7) << 12 | (ac0 & 7) << 9 | (a1 & 0xF) << 5 | (c1 & 0x1F)
    w1 = (b0 & 0xF) << 28 | (b1 & 0xF) << 24 | (aa1 & 7) << 21 | (ac1 & 7) << 18 | (d0 & 7) << 15
    w1 |= (ab0 & 7) << 12 | (ad0 & 7) << 9 | (d1 & 7) << 6 | (ab1 & 7) << 3 | (ad1 & 7)
    return [_u32(w0), _u32(w1)]


# Combiner selector values (rdp_common.hpp:83-174).
RGB_COMBINED, RGB_TEXEL0, RGB_TEXEL1, RGB_PRIM, RGB_SHADE, RGB_ENV = 0, 1, 2, 3, 4, 5
MULADD_ONE, MULADD_NOISE, MULADD_ZERO = 6, 7, 8
MULSUB_KEY_CENTER, MULSUB_K4, MULSUB_ZERO = 6, 7, 8
MUL_KEY_SCALE, MUL_COMBINED_ALPHA, MUL_TEXEL0_ALPHA, MUL_TEXEL1_ALPHA = 6, 7, 8, 9
MUL_PRIM_ALPHA, MUL_SHADE_ALPHA, MUL_ENV_ALPHA, MUL_LOD_FRAC, MUL_PRIM_LOD_FRAC, MUL_K5, MUL_ZERO = \
    10, 11, 12, 13, 14, 15, 16
ADD_ONE, ADD_ZERO = 6, 7
A_COMBINED, A_TEXEL0, A_TEXEL1, A_PRIM, A_SHADE, A_ENV, A_ONE, A_ZERO = range(8)
AMUL_LOD_FRAC, AMUL_PRIM_LOD_FRAC, AMUL_ZERO = 0, 6, 7


class CommandList:
    """Accumulates raw RDP command words."""

    def __init__(self):
        self._chunks: List[np.ndarray] = []
        self._pending: List[int] = []

    def _flush_pending(self):
        if self._pending:
            self._chunks.append(np.asarray(self._pending, dtype=np.uint32))
            self._pending = []

    def raw(self, words: Sequence[int]):
        self._pending.extend(_u32(w) for w in words)

    def extend_array(self, arr: np.ndarray):
        self._flush_pending()
        self._chunks.append(np.ascontiguousarray(arr, dtype=np.uint32).ravel())

    def words(self) -> np.ndarray:
        self._flush_pending()
        if not self._chunks:
            return np.zeros(0, dtype=np.uint32)
        return np.concatenate(self._chunks)

    # --- state commands ----------------------------------------------------
    def set_color_image(self, addr: int, width: int, size: int = SIZ_16, fmt: int = FMT_RGBA):
        self.raw([OP_SET_COLOR_IMAGE << 24 | (fmt & 7) << 21 | (size & 3) << 19 | ((width - 1) & 1023),
                  addr & 0xFFFFFF])

    def set_mask_image(self, addr: int):
        self.raw([OP_SET_MASK_IMAGE << 24, addr & 0xFFFFFF])

    def set_scissor(self, xlo: int, ylo: int, xhi: int, yhi: int, field: bool = False, keep_odd: bool = False):
        """Edges in 10.2 fixed point."""
        w1 = (xhi & 0xFFF) << 12 | (yhi & 0xFFF)
        if field:
            w1 |= 1 << 25
        if keep_odd:
            w1 |= 1 << 24
        self.raw([OP_SET_SCISSOR << 24 | (xlo & 0xFFF) << 12 | (ylo & 0xFFF), w1])

    def set_other_modes(self, modes: OtherModes):
        self.raw(modes.words())

    def set_combine(self, c0_rgb, c0_alpha, c1_rgb=None, c1_alpha=None):
        if c1_rgb is None:
            c1_rgb, c1_alpha = c0_rgb, c0_alpha
        self.raw(combine_words(c0_rgb, c0_alpha, c1_rgb, c1_alpha))

    def set_fill_color(self, color: int):
        self.raw([OP_SET_FILL_COLOR << 24, color])

    def set_fog_color(self, color: int):
        self.raw([OP_SET_FOG_COLOR << 24, color])

    def set_blend_color(self, color: int):
        self.raw([OP_SET_BLEND_COLOR << 24, color])

    def set_prim_color(self, color: int, min_level: int = 0, lod_frac: int = 0):
        self.raw([OP_SET_PRIM_COLOR << 24 | (min_level & 31) << 8 | (lod_frac & 0xFF), color])

    def set_env_color(self, color: int):
        self.raw([OP_SET_ENV_COLOR << 24, color])

    def set_prim_depth(self, z: int, dz: int):
        self.raw([OP_SET_PRIM_DEPTH << 24, (z & 0xFFFF) << 16 | (dz & 0xFFFF)])

    def set_convert(self, k: Sequence[int]):
        m = 0
        for i, v in enumerate(k):
            m |= (v & 0x1FF) << (45 - 9 * i)
        self.raw([OP_SET_CONVERT << 24 | (m >> 32) & 0xFFFFFF, m & 0xFFFFFFFF])

    def set_quirks(self, native_resolution_tex_rect: bool = True, native_texture_lod: bool = False):
        """MetaSetQuirks (rdp_device.cpp:948-968): byte 0 and byte 1 of the
        Quirks word.  Both matter only to upscaled rendering."""
        self.raw([0x04 << 24, int(native_resolution_tex_rect) | int(native_texture_lod) << 8])

    def sync_full(self):
        self.raw([OP_SYNC_FULL << 24, 0])

    def sync_pipe(self):
        self.raw([OP_SYNC_PIPE << 24, 0])

    def sync_load(self):
        self.raw([OP_SYNC_LOAD << 24, 0])

    def sync_tile(self):
        self.raw([OP_SYNC_TILE << 24, 0])

    # --- texture commands --------------------------------------------------
    def set_texture_image(self, addr: int, width: int, size: int, fmt: int = FMT_RGBA):
        self.raw([OP_SET_TEXTURE_IMAGE << 24 | (fmt & 7) << 21 | (size & 3) << 19 | ((width - 1) & 0x3FF),
                  addr & 0xFFFFFF])

    def set_tile(self, tile: int, fmt: int, size: int, tmem_offset: int, stride: int, palette: int = 0,
                 mask_s: int = 0, shift_s: int = 0, mask_t: int = 0, shift_t: int = 0,
                 clamp_s: bool = False, mirror_s: bool = False, clamp_t: bool = False, mirror_t: bool = False):
        """tmem_offset and stride in bytes (multiples of 8)."""
        w0 = OP_SET_TILE << 24 | (fmt & 7) << 21 | (size & 3) << 19 | ((stride >> 3) & 511) << 9 | ((tmem_offset >> 3) & 511)
        w1 = (tile & 7) << 24 | (palette & 15) << 20 | int(clamp_t) << 19 | int(mirror_t) << 18
        w1 |= (mask_t & 15) << 14 | (shift_t & 15) << 10 | int(clamp_s) << 9 | int(mirror_s) << 8
        w1 |= (mask_s & 15) << 4 | (shift_s & 15)
        self.raw([w0, w1])

    def _tile_cmd(self, op: int, tile: int, s0: int, t0: int, s1: int, t1: int):
        self.raw([op << 24 | (s0 & 0xFFF) << 12 | (t0 & 0xFFF), (tile & 7) << 24 | (s1 & 0xFFF) << 12 | (t1 & 0xFFF)])

    def load_tile(self, tile: int, s0: int, t0: int, s1: int, t1: int):
        """Coordinates in 10.2 fixed point."""
        self._tile_cmd(OP_LOAD_TILE, tile, s0, t0, s1, t1)

    def load_block(self, tile: int, s0: int, t0: int, s1: int, dxt: int):
        self._tile_cmd(OP_LOAD_BLOCK, tile, s0, t0, s1, dxt)

    def load_tlut(self, tile: int, s0: int, t0: int, s1: int, t1: int):
        self._tile_cmd(OP_LOAD_TLUT, tile, s0, t0, s1, t1)

    def set_tile_size(self, tile: int, s0: int, t0: int, s1: int, t1: int):
        self._tile_cmd(OP_SET_TILE_SIZE, tile, s0, t0, s1, t1)

    # --- primitives --------------------------------------------------------
    def fill_rectangle(self, x0: int, y0: int, x1: int, y1: int):
        """Upper-left (x0, y0) and lower-right (x1, y1) in 10.2 fixed point."""
        self.raw([OP_FILL_RECTANGLE << 24 | (x1 & 0xFFF) << 12 | (y1 & 0xFFF), (x0 & 0xFFF) << 12 | (y0 & 0xFFF)])

    def texture_rectangle(self, tile: int, x0: int, y0: int, x1: int, y1: int, s: int, t: int, dsdx: int, dtdy: int,
                          flip: bool = False):
        op = OP_TEXTURE_RECTANGLE_FLIP if flip else OP_TEXTURE_RECTANGLE
        self.raw([op << 24 | (x1 & 0xFFF) << 12 | (y1 & 0xFFF), (tile & 7) << 24 | (x0 & 0xFFF) << 12 | (y0 & 0xFFF),
                  (s & 0xFFFF) << 16 | (t & 0xFFFF), (dsdx & 0xFFFF) << 16 | (dtdy & 0xFFFF)])


# ----------------------------------------------------------------------------
# Triangle setup
# ----------------------------------------------------------------------------
def _round_away_div(num: np.ndarray, den: np.ndarray) -> np.ndarray:
    """Integer division rounding away from zero (triangle_converter.cpp:121-130)."""
    mag = (np.abs(num) + den - 1) // den
    return np.where(num < 0, -mag, mag)


def _pack_pair(a: np.ndarray, b: np.ndarray):
    """Two 16.16 values -> (integer-halves word, fraction-halves word)."""
    a = a.astype(np.int64) & 0xFFFFFFFF
    b = b.astype(np.int64) & 0xFFFFFFFF
    hi = (a & 0xFFFF0000) | ((b >> 16) & 0xFFFF)
    lo = ((a << 16) & 0xFFFF0000) | (b & 0xFFFF)
    return hi, lo


def triangle_commands(xy_q: np.ndarray, rgba: Optional[np.ndarray] = None, z: Optional[np.ndarray] = None,
                      uv: Optional[np.ndarray] = None, w: Optional[np.ndarray] = None, perspective: bool = False,
                      tile: int = 0, max_level: int = 0) -> np.ndarray:
    """Build triangle commands for a batch.

    xy_q:  (N, 3, 2) integer vertex positions in quarter pixels.
    rgba:  (N, 3, 4) float colour in [0, 255] (shade triangles).
    z:     (N, 3) float depth in [0, 0x3ffff] (z-buffered triangles).
    uv:    (N, 3, 2) float texel coordinates (textured triangles).
    w:     (N, 3) float clip w > 0, only with perspective=True.

    Returns a (M, L) uint32 array, one command per row, where degenerate
    (zero-area) triangles have been dropped.
    """
    xy_q = np.asarray(xy_q, dtype=np.int64)
    n = xy_q.shape[0]
    shade = rgba is not None
    zbuf = z is not None
    tex = uv is not None
    op = OP_FILL_TRIANGLE | (4 if shade else 0) | (2 if tex else 0) | (1 if zbuf else 0)
    length = int(COMMAND_LENGTHS[op])

    xs = xy_q[:, :, 0]
    ys = xy_q[:, :, 1]
    # Sort vertices by Y, tie-break on X (triangle_converter.cpp:145-158).
    key = ys * 65536 + (xs + 32768)
    order = np.argsort(key, axis=1, kind="stable")
    ar = np.arange(n)[:, None]
    xs = xs[ar, order]
    ys = ys[ar, order]
    xa, xb, xc = xs[:, 0], xs[:, 1], xs[:, 2]
    y_lo, y_mid, y_hi = ys[:, 0], ys[:, 1], ys[:, 2]

    # Signed area on sorted vertices; zero area is dropped.
    ab_x, bc_x, ca_x = xb - xa, xc - xb, xa - xc
    ab_y, bc_y, ca_y = y_mid - y_lo, y_hi - y_mid, y_lo - y_hi
    area = ab_x * bc_y - ab_y * bc_x
    keep = area != 0

    x_a = xa << 14
    x_b = xa << 14
    x_c = xb << 14
    dxdy_a = _round_away_div((xc - xa) << 16, np.maximum(1, y_hi - y_lo)) & ~7
    dxdy_b = _round_away_div((xb - xa) << 16, np.maximum(1, y_mid - y_lo)) & ~7
    dxdy_c = _round_away_div((xc - xb) << 16, np.maximum(1, y_hi - y_mid)) & ~7
    sub = y_lo & 3
    x_a = x_a - (dxdy_a >> 2) * sub
    x_b = x_b - (dxdy_b >> 2) * sub
    right_major = dxdy_b < dxdy_a

    words = np.zeros((n, length), dtype=np.int64)
    words[:, 0] = (op << 24) | (np.where(right_major, 0, 1) << 23) | ((max_level & 7) << 19) | ((tile & 7) << 16) | (y_hi & 0x3FFF)
    words[:, 1] = ((y_mid & 0x3FFF) << 16) | (y_lo & 0x3FFF)
    words[:, 2] = x_c & 0x0FFFFFFF
    words[:, 3] = dxdy_c & 0x3FFFFFFF
    words[:, 4] = x_a & 0x0FFFFFFF
    words[:, 5] = dxdy_a & 0xFFFFFFFF
    words[:, 6] = x_b & 0x0FFFFFFF
    words[:, 7] = dxdy_b & 0x3FFFFFFF

    safe_area = np.where(keep, area, 1).astype(np.float64)
    inv_area = 4.0 / safe_area
    dxdy = dxdy_a.astype(np.float64) / 65536.0
    yfrac = (y_lo & 3).astype(np.float64) / 4.0

    def plane(vals: np.ndarray, scale: float):
        """vals: (N, 3) unsorted per-vertex values -> start, dx, de, dy (int64)."""
        v = vals[ar, order].astype(np.float64)
        va, vb, vc = v[:, 0], v[:, 1], v[:, 2]
        ddx = -inv_area * (ab_y * vc + ca_y * vb + bc_y * va)
        ddy = inv_area * (ab_x * vc + ca_x * vb + bc_x * va)
        dde = ddy + ddx * dxdy
        start = va - yfrac * dde
        q = lambda a: np.round(a * scale).astype(np.int64)
        return q(start), q(ddx), q(dde), q(ddy)

    base = 8
    if shade:
        comps = [plane(np.asarray(rgba, dtype=np.float64)[:, :, c], 65536.0) for c in range(4)]
        # comps[c] = (start, dx, de, dy)
        for k, (blk_hi, blk_lo) in enumerate(((0, 4), (2, 6), (8, 12), (10, 14))):
            r_hi, r_lo = _pack_pair(comps[0][k], comps[1][k])
            b_hi, b_lo = _pack_pair(comps[2][k], comps[3][k])
            words[:, base + blk_hi] = r_hi
            words[:, base + blk_hi + 1] = b_hi
            words[:, base + blk_lo] = r_lo
            words[:, base + blk_lo + 1] = b_lo
        base += 16
    if tex:
        uv = np.asarray(uv, dtype=np.float64)
        if perspective:
            ww = np.asarray(w, dtype=np.float64)
            min_w = ww.min(axis=1, keepdims=True) * 0.49
            iw = min_w / ww
            s_vals = uv[:, :, 0] * iw * 64.0
            t_vals = uv[:, :, 1] * iw * 64.0
            w_vals = iw
            s_p = plane(s_vals, 65536.0)
            t_p = plane(t_vals, 65536.0)
            w_p = plane(w_vals, 4294967296.0)
        else:
            s_p = plane(uv[:, :, 0] * 32.0, 65536.0)
            t_p = plane(uv[:, :, 1] * 32.0, 65536.0)
            w_p = tuple(np.zeros(n, dtype=np.int64) for _ in range(4))
        for k, (blk_hi, blk_lo) in enumerate(((0, 4), (2, 6), (8, 12), (10, 14))):
            st_hi, st_lo = _pack_pair(s_p[k], t_p[k])
            w_hi, w_lo = _pack_pair(w_p[k], np.zeros(n, dtype=np.int64))
            words[:, base + blk_hi] = st_hi
            words[:, base + blk_hi + 1] = w_hi
            words[:, base + blk_lo] = st_lo
            words[:, base + blk_lo + 1] = w_lo
        base += 16
    if zbuf:
        zp = plane(np.asarray(z, dtype=np.float64), 8192.0)
        for k in range(4):
            words[:, base + k] = zp[k] & 0xFFFFFFFF
        base += 4

    return (words[keep] & 0xFFFFFFFF).astype(np.uint32)


# ----------------------------------------------------------------------------
# RDRAM helpers: the N64 address space stored as host-endian 32-bit words.
# ----------------------------------------------------------------------------
def rdram_write16(rdram: np.ndarray, addr: int, values: np.ndarray):
    """Write big-endian halfwords starting at byte address addr (2-aligned)."""
    h = rdram.view(np.uint16)
    idx = (addr >> 1) + np.arange(len(values))
    h[idx ^ 1] = values.astype(np.uint16)


def rdram_write8(rdram: np.ndarray, addr: int, values: np.ndarray):
    b = rdram.view(np.uint8)
    idx = addr + np.arange(len(values))
    b[idx ^ 3] = values.astype(np.uint8)


def rdram_read16(rdram: np.ndarray, addr: int, count: int) -> np.ndarray:
    h = rdram.view(np.uint16)
    idx = (addr >> 1) + np.arange(count)
    return h[idx ^ 1]


# ----------------------------------------------------------------------------
# Workloads
# ----------------------------------------------------------------------------
@dataclasses.dataclass
class Frame:
    """One frame of a stream: RDP words plus the VI register state for scanout."""
    words: np.ndarray
    vi: Optional[Dict[int, int]] = None


@dataclasses.dataclass
class Stream:
    """A seeded command stream with its initial RDRAM contents."""
    name: str
    seed: int
    frames: List[Frame]
    rdram_init: Optional[np.ndarray] = None      # uint32 words, len RDRAM_SIZE/4 (None = zero)
    hidden_init: int = 0                          # fill value for hidden RDRAM
    color_addr: int = 0
    depth_addr: int = 0
    width: int = 0
    height: int = 0
    color_size: int = SIZ_16
    num_triangles: int = 0

    def all_words(self) -> np.ndarray:
        return np.concatenate([f.words for f in self.frames]) if self.frames else np.zeros(0, np.uint32)


COLOR_ADDR = 0x100000
DEPTH_ADDR = 0x300000
TEXTURE_ADDR = 0x500000


def _clear_passes(cl: CommandList, width: int, height: int, color_size: int, clear_color: bool,
                  color_addr: int = COLOR_ADDR, depth_addr: int = DEPTH_ADDR):
    """Standard frame start: fill the Z buffer with the far value, optionally
    fill the colour buffer, then bind colour + depth for drawing."""
    fill = OtherModes(cycle_type=3)
    cl.set_other_modes(fill)
    cl.set_scissor(0, 0, width << 2, height << 2)
    cl.set_color_image(depth_addr, width, SIZ_16)
    cl.set_fill_color(0xFFFCFFFC)
    cl.fill_rectangle(0, 0, (width - 1) << 2, (height - 1) << 2)
    cl.sync_pipe()
    cl.set_color_image(color_addr, width, color_size)
    cl.set_mask_image(depth_addr)
    if clear_color:
        cl.set_fill_color(0x00010001 if color_size == SIZ_16 else 0x000000FF)
        cl.fill_rectangle(0, 0, (width - 1) << 2, (height - 1) << 2)
        cl.sync_pipe()


def config0_stream(seed: int = 0x52445000, num_triangles: int = 256, width: int = 320, height: int = 240) -> Stream:
    """Plain method: 256 Gouraud triangles, 1-cycle SHADE combiner, Z compare
    and update, RGBA5551 colour + 16-bit Z, one frame."""
    rng = np.random.default_rng(seed)
    cl = CommandList()
    _clear_passes(cl, width, height, SIZ_16, clear_color=False)
    modes = OtherModes(cycle_type=0, depth_test=True, depth_write=True, z_mode=0,
                       blend=((0, 0, 1, 0), (0, 0, 1, 0)))
    cl.set_other_modes(modes)
    shade = ((MULADD_ZERO, MULSUB_ZERO, MUL_ZERO, RGB_SHADE), (A_ZERO, A_ZERO, AMUL_ZERO, A_SHADE))
    cl.set_combine(shade[0], shade[1])
    xy = np.stack([rng.integers(0, width * 4, size=(num_triangles, 3)),
                   rng.integers(0, height * 4, size=(num_triangles, 3))], axis=-1)
    rgba = rng.integers(0, 256, size=(num_triangles, 3, 4)).astype(np.float64)
    z = rng.integers(0, 0x40000, size=(num_triangles, 3)).astype(np.float64)
    tris = triangle_commands(xy, rgba=rgba, z=z)
    cl.extend_array(tris)
    cl.sync_full()
    return Stream(name="c0", seed=seed, frames=[Frame(cl.words())], hidden_init=0, color_addr=COLOR_ADDR,
                  depth_addr=DEPTH_ADDR, width=width, height=height, color_size=SIZ_16,
                  num_triangles=int(tris.shape[0]))


def _log_uniform_triangles(rng, n: int, width: int, height: int, lo: float, hi: float):
    """Triangles whose extent is log-uniform in [lo, hi] pixels at a uniform
    position; vertices in quarter pixels, clamped to the guard band."""
    size = np.exp(rng.uniform(math.log(lo), math.log(hi), size=n))
    cx = rng.uniform(0, width, size=n)
    cy = rng.uniform(0, height, size=n)
    ang = rng.uniform(0, 2 * math.pi, size=(n, 3))
    rad = size[:, None] * rng.uniform(0.35, 0.6, size=(n, 3))
    x = cx[:, None] + rad * np.cos(ang)
    y = cy[:, None] + rad * np.sin(ang)
    xq = np.clip(np.round(x * 4), -1024 * 4, 1023 * 4).astype(np.int64)
    yq = np.clip(np.round(y * 4), 0, (height - 1) * 4 + 3).astype(np.int64)
    return np.stack([xq, yq], axis=-1)


def _random_texture_rgba16(rng, tex_w: int, tex_h: int) -> np.ndarray:
    return rng.integers(0, 1 << 16, size=tex_w * tex_h, dtype=np.int64).astype(np.uint16)


def config1_stream(seed: int = 0x52446000, frames: int = 60, tris_per_frame: int = 4096, width: int = 640,
                   height: int = 480, num_textures: int = 8, tris_per_texture: int = 64) -> Stream:
    """Textured 1-cycle pass: RGBA8888 colour + Z, 32x32 RGBA16 textures loaded
    with LoadTile, bilinear + perspective, TEX0*SHADE, alpha blend, Z on."""
    rng = np.random.default_rng(seed)
    rdram = np.zeros(RDRAM_SIZE // 4, dtype=np.uint32)
    tex_w = tex_h = 32
    tex_bytes = tex_w * tex_h * 2
    for i in range(num_textures):
        rdram_write16(rdram, TEXTURE_ADDR + i * tex_bytes, _random_texture_rgba16(rng, tex_w, tex_h))

    out_frames = []
    total = 0
    shade_tex = ((RGB_TEXEL0, MULSUB_ZERO, RGB_SHADE, ADD_ZERO), (A_TEXEL0, A_ZERO, A_SHADE, A_ZERO))
    modes = OtherModes(cycle_type=0, perspective=True, sample_quad=True, bilerp0=True, bilerp1=True,
                       depth_test=True, depth_write=True, z_mode=0, force_blend=True,
                       blend=((0, 0, 1, 0), (0, 0, 1, 0)))
    for f in range(frames):
        cl = CommandList()
        _clear_passes(cl, width, height, SIZ_32, clear_color=True)
        cl.set_other_modes(modes)
        cl.set_combine(shade_tex[0], shade_tex[1])
        n = tris_per_frame
        xy = _log_uniform_triangles(rng, n, width, height, 2.0, 200.0)
        rgba = rng.integers(0, 256, size=(n, 3, 4)).astype(np.float64)
        z = rng.uniform(0, 0x3FFFF, size=(n, 3))
        uv = rng.uniform(-8.0, 40.0, size=(n, 3, 2))
        w = rng.uniform(0.3, 1.0, size=(n, 3))
        tris = triangle_commands(xy, rgba=rgba, z=z, uv=uv, w=w, perspective=True, tile=0, max_level=0)
        total += tris.shape[0]
        for start in range(0, tris.shape[0], tris_per_texture):
            t = int(rng.integers(0, num_textures))
            cl.set_texture_image(TEXTURE_ADDR + t * tex_bytes, tex_w, SIZ_16, FMT_RGBA)
            cl.set_tile(7, FMT_RGBA, SIZ_16, 0, tex_w * 2)
            cl.sync_load()
            cl.load_tile(7, 0, 0, (tex_w - 1) << 2, (tex_h - 1) << 2)
            cl.sync_tile()
            cl.set_tile(0, FMT_RGBA, SIZ_16, 0, tex_w * 2, mask_s=5, mask_t=5)
            cl.set_tile_size(0, 0, 0, (tex_w - 1) << 2, (tex_h - 1) << 2)
            cl.extend_array(tris[start:start + tris_per_texture])
        cl.sync_full()
        out_frames.append(Frame(cl.words()))
    return Stream(name="c1", seed=seed, frames=out_frames, rdram_init=rdram, hidden_init=0, color_addr=COLOR_ADDR,
                  depth_addr=DEPTH_ADDR, width=width, height=height, color_size=SIZ_32, num_triangles=total)


# ----------------------------------------------------------------------------
# 2-cycle stress workload
# ----------------------------------------------------------------------------
VI_TYPE_RGBA5551 = 2
VI_AA_RESAMP_EXTRA = 1 << 8
VI_DIVOT = 1 << 4
VI_DITHER_FILTER = 1 << 16


def vi_start_register(start: int, end: int) -> int:
    """make_vi_start_register, rdp_common.hpp:365-368."""
    return ((start & 0x3FF) << 16) | (end & 0x3FF)


def vi_scale_register(scale: int, bias: int) -> int:
    """make_vi_scale_register, rdp_common.hpp:370-373."""
    return ((bias & 0xFFF) << 16) | (scale & 0xFFF)


MIP_SIZES = (32, 16, 8, 4)
TEXTURE_STRIDE = 0x2000       # RDRAM bytes per CI texture (4 levels + TLUT)
TLUT_OFFSET = 0x1800


def _ci_level_layout(size_bits: int):
    """TMEM byte offset and line stride of each mip level (lower TMEM)."""
    offs, strides, off = [], [], 0
    for s in MIP_SIZES:
        line = max(8, (s * (8 if size_bits == SIZ_8 else 4) // 8 + 7) & ~7)
        offs.append(off)
        strides.append(line)
        off += line * s
    return offs, strides


def _write_ci_texture(rdram: np.ndarray, rng, index: int, size_bits: int) -> int:
    """Random CI4/CI8 mip chain and its TLUT (RGBA5551 entries) in RDRAM."""
    base = TEXTURE_ADDR + index * TEXTURE_STRIDE
    addr = base
    for s in MIP_SIZES:
        nbytes = s * s if size_bits == SIZ_8 else s * s // 2
        rdram_write8(rdram, addr, rng.integers(0, 256, size=nbytes))
        addr += (nbytes + 7) & ~7
    entries = 256 if size_bits == SIZ_8 else 256  # CI4 uses 16 palettes of 16
    rdram_write16(rdram, base + TLUT_OFFSET, rng.integers(0, 1 << 16, size=entries))
    return base


def _emit_ci_texture(cl: CommandList, index: int, size_bits: int, palette: int):
    """Load the 4 mip levels with LoadTile (CI4 as 8-bit texels of half the
    width) and the TLUT with LoadTLUT, then set render tiles 0-3."""
    base = TEXTURE_ADDR + index * TEXTURE_STRIDE
    offs, strides = _ci_level_layout(size_bits)
    addr = base
    cl.sync_load()
    for lvl, s in enumerate(MIP_SIZES):
        nbytes = s * s if size_bits == SIZ_8 else s * s // 2
        w = s if size_bits == SIZ_8 else max(s // 2, 1)
        cl.set_texture_image(addr, w, SIZ_8, FMT_CI)
        cl.set_tile(7, FMT_CI, SIZ_8, offs[lvl], strides[lvl])
        cl.load_tile(7, 0, 0, (w - 1) << 2, (s - 1) << 2)
        addr += (nbytes + 7) & ~7
    n = 256 if size_bits == SIZ_8 else 16
    first = 0 if size_bits == SIZ_8 else palette * 16
    cl.set_texture_image(base + TLUT_OFFSET + 2 * first, n, SIZ_16, FMT_RGBA)
    cl.set_tile(7, FMT_RGBA, SIZ_4, 0x800 + first * 8, 0)
    cl.load_tlut(7, 0, 0, (n - 1) << 2, 0)
    cl.sync_tile()
    for lvl, s in enumerate(MIP_SIZES):
        m = s.bit_length() - 1
        cl.set_tile(lvl, FMT_CI, size_bits, offs[lvl], strides[lvl], palette=palette if size_bits == SIZ_4 else 0,
                    mask_s=m, mask_t=m, shift_s=lvl, shift_t=lvl)
        cl.set_tile_size(lvl, 0, 0, (s - 1) << 2, (s - 1) << 2)


C2_MODES = OtherModes(cycle_type=1, perspective=True, tex_lod=True, tlut=True, sample_quad=True, bilerp0=True,
                      bilerp1=True, blend=((3, 2, 0, 0), (0, 0, 1, 1)), aa=True, image_read=True,
                      alpha_cvg_select=True, depth_test=True, depth_write=True, z_mode=0, coverage_mode=0)
# cycle 0: lerp(TEXEL0, TEXEL1, LOD_FRAC); cycle 1: COMBINED * SHADE
C2_COMBINE = (((RGB_TEXEL1, RGB_TEXEL0, MUL_LOD_FRAC, RGB_TEXEL0), (A_TEXEL1, A_TEXEL0, AMUL_LOD_FRAC, A_TEXEL0)),
              ((RGB_COMBINED, MULSUB_ZERO, RGB_SHADE, ADD_ZERO), (A_COMBINED, A_ZERO, A_SHADE, A_ZERO)))
# Sprites (texture rectangles): no depth, 2-cycle, cycle 1 passes the texture
# through scaled by the environment colour.
C2_RECT_MODES = OtherModes(cycle_type=1, tlut=True, sample_quad=True, bilerp0=True, bilerp1=True,
                           blend=((0, 0, 0, 0), (0, 0, 1, 1)), aa=True, image_read=True, alpha_cvg_select=True,
                           coverage_mode=0)
C2_RECT_COMBINE = (((MULADD_ZERO, MULSUB_ZERO, MUL_ZERO, RGB_TEXEL0), (A_ZERO, A_ZERO, AMUL_ZERO, A_TEXEL0)),
                   ((RGB_COMBINED, MULSUB_ZERO, RGB_ENV, ADD_ZERO), (A_ZERO, A_ZERO, AMUL_ZERO, A_COMBINED)))


def c2_vi_registers(width: int, height: int, origin: int = COLOR_ADDR) -> Dict[int, int]:
    """RGBA5551 scanout with AA (resample extra), divot and dither filter."""
    return {VI_CONTROL: VI_TYPE_RGBA5551 | VI_AA_RESAMP_EXTRA | VI_DIVOT | VI_DITHER_FILTER,
            VI_ORIGIN: origin, VI_WIDTH: width, VI_V_SYNC: 525,
            VI_V_START: vi_start_register(34, 34 + 480), VI_H_START: vi_start_register(108, 108 + 640),
            VI_X_SCALE: vi_scale_register((width * 1024) // 640, 0),
            VI_Y_SCALE: vi_scale_register((height * 1024) // 240, 0)}


def _emit_c2(cl: CommandList, rng, n: int, width: int, height: int, num_textures: int,
             prims_per_texture: int = 64) -> int:
    """2-cycle CI textured triangles with mip LOD, fog and coverage AA, plus
    20 % texture rectangles drawn as sprites; returns the number of
    primitives emitted."""
    cl.set_fog_color(int(rng.integers(0, 1 << 32, dtype=np.int64)))
    cl.set_env_color(int(rng.integers(0x40404000, 1 << 32, dtype=np.int64)))
    n_rect = n // 5
    n_tri = n - n_rect
    xy = _log_uniform_triangles(rng, n_tri, width, height, 4.0, 160.0)
    rgba = rng.integers(32, 256, size=(n_tri, 3, 4)).astype(np.float64)
    z = rng.uniform(0, 0x3FFFF, size=(n_tri, 3))
    uv = rng.uniform(-64.0, 256.0, size=(n_tri, 3, 2))
    w = rng.uniform(0.3, 1.0, size=(n_tri, 3))
    tris = triangle_commands(xy, rgba=rgba, z=z, uv=uv, w=w, perspective=True, tile=0, max_level=3)
    rects = []
    for _ in range(n_rect):
        x0 = int(rng.integers(0, width * 4))
        y0 = int(rng.integers(0, height * 4))
        rw, rh = int(rng.integers(8, 48)) * 4, int(rng.integers(8, 48)) * 4
        rects.append((x0, y0, x0 + rw, y0 + rh, int(rng.integers(-512, 2048)), int(rng.integers(-512, 2048)),
                      int(rng.choice([512, 1024, 2048])), int(rng.choice([512, 1024, 2048]))))
    ti = ri = 0
    emitted = 0
    per_group_rects = max(prims_per_texture // 5, 1)
    while ti < tris.shape[0] or ri < len(rects):
        t = int(rng.integers(0, num_textures))
        size_bits = SIZ_8 if t % 2 == 0 else SIZ_4
        _emit_ci_texture(cl, t, size_bits, palette=int(rng.integers(0, 16)))
        k = min(prims_per_texture - per_group_rects, tris.shape[0] - ti)
        if k > 0:
            cl.sync_pipe()
            cl.set_other_modes(C2_MODES)
            cl.set_combine(C2_COMBINE[0][0], C2_COMBINE[0][1], C2_COMBINE[1][0], C2_COMBINE[1][1])
            cl.extend_array(tris[ti:ti + k])
            ti += k
            emitted += k
        k = min(per_group_rects, len(rects) - ri)
        if k > 0:
            cl.sync_pipe()
            cl.set_other_modes(C2_RECT_MODES)
            cl.set_combine(C2_RECT_COMBINE[0][0], C2_RECT_COMBINE[0][1], C2_RECT_COMBINE[1][0],
                           C2_RECT_COMBINE[1][1])
            for x0, y0, x1, y1, s0, t0, dsdx, dtdy in rects[ri:ri + k]:
                cl.texture_rectangle(0, x0, y0, x1, y1, s0, t0, dsdx, dtdy)
            ri += k
            emitted += k
    return emitted


def config2_stream(seed: int = 0x52447000, frames: int = 30, prims_per_frame: int = 2560, width: int = 640,
                   height: int = 480, num_textures: int = 8) -> Stream:
    """2-cycle stress pass: 640x480 RGBA5551 + Z, CI4/CI8 textures with random
    16/256-entry TLUTs, mip LOD over 4 levels, fog blender in cycle 0 and
    coverage AA in cycle 1, 20 % texture rectangles; every frame is scanned
    out through the VI with AA, divot and dither filters."""
    rng = np.random.default_rng(seed)
    rdram = np.zeros(RDRAM_SIZE // 4, dtype=np.uint32)
    for t in range(num_textures):
        _write_ci_texture(rdram, rng, t, SIZ_8 if t % 2 == 0 else SIZ_4)
    out_frames = []
    total = 0
    for f in range(frames):
        cl = CommandList()
        _clear_passes(cl, width, height, SIZ_16, clear_color=True)
        total += _emit_c2(cl, rng, prims_per_frame, width, height, num_textures)
        cl.sync_full()
        out_frames.append(Frame(cl.words(), vi=c2_vi_registers(width, height)))
    return Stream(name="c2", seed=seed, frames=out_frames, rdram_init=rdram, hidden_init=0, color_addr=COLOR_ADDR,
                  depth_addr=DEPTH_ADDR, width=width, height=height, color_size=SIZ_16, num_triangles=total)


# ----------------------------------------------------------------------------
# Mixed batch (64 independent streams)
# ----------------------------------------------------------------------------
SHADE_COMBINE = ((MULADD_ZERO, MULSUB_ZERO, MUL_ZERO, RGB_SHADE), (A_ZERO, A_ZERO, AMUL_ZERO, A_SHADE))
TEX_SHADE_COMBINE = ((RGB_TEXEL0, MULSUB_ZERO, RGB_SHADE, ADD_ZERO), (A_TEXEL0, A_ZERO, A_SHADE, A_ZERO))


def _emit_shade_z(cl: CommandList, rng, n: int, width: int, height: int, uniform_vertices: bool) -> int:
    cl.set_other_modes(OtherModes(cycle_type=0, depth_test=True, depth_write=True,
                                  blend=((0, 0, 1, 0), (0, 0, 1, 0))))
    cl.set_combine(*SHADE_COMBINE)
    if uniform_vertices:
        xy = np.stack([rng.integers(0, width * 4, size=(n, 3)), rng.integers(0, height * 4, size=(n, 3))], axis=-1)
    else:
        xy = _log_uniform_triangles(rng, n, width, height, 2.0, 200.0)
    rgba = rng.integers(0, 256, size=(n, 3, 4)).astype(np.float64)
    z = rng.integers(0, 0x40000, size=(n, 3)).astype(np.float64)
    tris = triangle_commands(xy, rgba=rgba, z=z)
    cl.extend_array(tris)
    return int(tris.shape[0])


def _emit_textured(cl: CommandList, rng, n: int, width: int, height: int, num_textures: int,
                   tris_per_texture: int = 64) -> int:
    tex_w = tex_h = 32
    tex_bytes = tex_w * tex_h * 2
    cl.set_other_modes(OtherModes(cycle_type=0, perspective=True, sample_quad=True, bilerp0=True, bilerp1=True,
                                  depth_test=True, depth_write=True, force_blend=True,
                                  blend=((0, 0, 1, 0), (0, 0, 1, 0))))
    cl.set_combine(*TEX_SHADE_COMBINE)
    xy = _log_uniform_triangles(rng, n, width, height, 2.0, 200.0)
    rgba = rng.integers(0, 256, size=(n, 3, 4)).astype(np.float64)
    z = rng.uniform(0, 0x3FFFF, size=(n, 3))
    uv = rng.uniform(-8.0, 40.0, size=(n, 3, 2))
    w = rng.uniform(0.3, 1.0, size=(n, 3))
    tris = triangle_commands(xy, rgba=rgba, z=z, uv=uv, w=w, perspective=True)
    for start in range(0, tris.shape[0], tris_per_texture):
        t = int(rng.integers(0, num_textures))
        cl.set_texture_image(TEXTURE_ADDR + t * tex_bytes, tex_w, SIZ_16, FMT_RGBA)
        cl.set_tile(7, FMT_RGBA, SIZ_16, 0, tex_w * 2)
        cl.sync_load()
        cl.load_tile(7, 0, 0, (tex_w - 1) << 2, (tex_h - 1) << 2)
        cl.sync_tile()
        cl.set_tile(0, FMT_RGBA, SIZ_16, 0, tex_w * 2, mask_s=5, mask_t=5)
        cl.set_tile_size(0, 0, 0, (tex_w - 1) << 2, (tex_h - 1) << 2)
        cl.extend_array(tris[start:start + tris_per_texture])
    return int(tris.shape[0])


BATCH_MODES = ("c0", "c1", "c2")


def batch_stream(index: int, seed_base: int = 0x52449000, modes: Sequence[str] = BATCH_MODES,
                 min_tris: int = 32, max_tris: int = 8192) -> Stream:
    """One stream of the mixed batch: log-uniform primitive count, 320x240 or
    640x480 at 16 or 32 bits per pixel, render mode drawn from the plain
    shaded (c0), textured (c1) and 2-cycle stress (c2) workloads."""
    seed = seed_base + index
    rng = np.random.default_rng(seed)
    n = int(round(math.exp(rng.uniform(math.log(min_tris), math.log(max_tris)))))
    width, height = ((320, 240), (640, 480))[int(rng.integers(0, 2))]
    size = (SIZ_16, SIZ_32)[int(rng.integers(0, 2))]
    mode = modes[int(rng.integers(0, len(modes)))]
    rdram = None
    cl = CommandList()
    _clear_passes(cl, width, height, size, clear_color=(mode != "c0"))
    if mode == "c1":
        rdram = np.zeros(RDRAM_SIZE // 4, dtype=np.uint32)
        num_textures = 8
        for i in range(num_textures):
            rdram_write16(rdram, TEXTURE_ADDR + i * 2048, _random_texture_rgba16(rng, 32, 32))
        tris = _emit_textured(cl, rng, n, width, height, num_textures)
    elif mode == "c2":
        rdram = np.zeros(RDRAM_SIZE // 4, dtype=np.uint32)
        for i in range(8):
            _write_ci_texture(rdram, rng, i, SIZ_8 if i % 2 == 0 else SIZ_4)
        tris = _emit_c2(cl, rng, n, width, height, 8)
    else:
        tris = _emit_shade_z(cl, rng, n, width, height, uniform_vertices=(n <= 512))
    cl.sync_full()
    return Stream(name=f"b{index}:{mode}:{width}x{height}x{4 << size}", seed=seed, frames=[Frame(cl.words())],
                  rdram_init=rdram, hidden_init=0, color_addr=COLOR_ADDR, depth_addr=DEPTH_ADDR, width=width,
                  height=height, color_size=size, num_triangles=tris)


DENSE_MODES = ("fill", "1cycle", "1cycle_tex", "2cycle")


def dense_stream(mode: str, width: int = 640, height: int = 480, layers: int = 16, seed: int = 0x5244D000) -> Stream:
    """Per-fragment cost calibration: frame 0 clears Z (and colour); frame 1
    draws `layers` primitives that each cover every pixel of the framebuffer,
    nearer one after the other, in one of the batch workloads' render modes
    (fill rectangles; the c0 shaded, c1 textured or c2 two-cycle states with
    one triangle per layer whose vertices lie outside the screen).  Every
    8x8 tile of frame 1 is fully covered by every layer, so the raster
    launch's instruction count over its fragments is the mode's cost per
    covered fragment with no idle lanes."""
    rng = np.random.default_rng(seed)
    size = SIZ_32 if mode == "1cycle_tex" else SIZ_16
    rdram = None
    cl = CommandList()
    _clear_passes(cl, width, height, size, clear_color=True)
    cl.sync_full()
    frames = [Frame(cl.words())]
    cl = CommandList()
    xy = np.tile(np.array([[0, 0], [width * 8, 0], [0, height * 8]], dtype=np.int64), (layers, 1, 1))
    z = np.repeat((0x30000 - 0x800 * np.arange(layers, dtype=np.float64))[:, None], 3, axis=1)
    rgba = rng.integers(32, 256, size=(layers, 3, 4)).astype(np.float64)
    if mode == "fill":
        cl.set_other_modes(OtherModes(cycle_type=3))
        for k in range(layers):
            cl.set_fill_color(int(rng.integers(0, 1 << 32, dtype=np.int64)))
            cl.fill_rectangle(0, 0, (width - 1) << 2, (height - 1) << 2)
        n = layers
    elif mode == "1cycle":
        cl.set_other_modes(OtherModes(cycle_type=0, depth_test=True, depth_write=True,
                                      blend=((0, 0, 1, 0), (0, 0, 1, 0))))
        cl.set_combine(*SHADE_COMBINE)
        tris = triangle_commands(xy, rgba=rgba, z=z)
        cl.extend_array(tris)
        n = int(tris.shape[0])
    elif mode == "1cycle_tex":
        rdram = np.zeros(RDRAM_SIZE // 4, dtype=np.uint32)
        rdram_write16(rdram, TEXTURE_ADDR, _random_texture_rgba16(rng, 32, 32))
        cl.set_other_modes(OtherModes(cycle_type=0, perspective=True, sample_quad=True, bilerp0=True, bilerp1=True,
                                      depth_test=True, depth_write=True, force_blend=True,
                                      blend=((0, 0, 1, 0), (0, 0, 1, 0))))
        cl.set_combine(*TEX_SHADE_COMBINE)
        cl.set_texture_image(TEXTURE_ADDR, 32, SIZ_16, FMT_RGBA)
        cl.set_tile(7, FMT_RGBA, SIZ_16, 0, 64)
        cl.sync_load()
        cl.load_tile(7, 0, 0, 31 << 2, 31 << 2)
        cl.sync_tile()
        cl.set_tile(0, FMT_RGBA, SIZ_16, 0, 64, mask_s=5, mask_t=5)
        cl.set_tile_size(0, 0, 0, 31 << 2, 31 << 2)
        uv = np.tile(np.array([[0.0, 0.0], [160.0, 0.0], [0.0, 120.0]]), (layers, 1, 1))
        w = np.tile(np.array([1.0, 0.6, 0.8]), (layers, 1))
        tris = triangle_commands(xy, rgba=rgba, z=z, uv=uv, w=w, perspective=True)
        cl.extend_array(tris)
        n = int(tris.shape[0])
    elif mode == "2cycle":
        rdram = np.zeros(RDRAM_SIZE // 4, dtype=np.uint32)
        _write_ci_texture(rdram, rng, 0, SIZ_8)
        cl.set_fog_color(0x80604020)
        cl.set_env_color(0x80808080)
        _emit_ci_texture(cl, 0, SIZ_8, palette=0)
        cl.sync_pipe()
        cl.set_other_modes(C2_MODES)
        cl.set_combine(C2_COMBINE[0][0], C2_COMBINE[0][1], C2_COMBINE[1][0], C2_COMBINE[1][1])
        uv = np.tile(np.array([[0.0, 0.0], [400.0, 0.0], [0.0, 300.0]]), (layers, 1, 1))
        w = np.tile(np.array([1.0, 0.6, 0.8]), (layers, 1))
        tris = triangle_commands(xy, rgba=rgba, z=z, uv=uv, w=w, perspective=True, tile=0, max_level=3)
        cl.extend_array(tris)
        n = int(tris.shape[0])
    else:
        raise ValueError(mode)
    cl.sync_full()
    frames.append(Frame(cl.words()))
    return Stream(name=f"dense:{mode}:{width}x{height}", seed=seed, frames=frames, rdram_init=rdram, hidden_init=0,
                  color_addr=COLOR_ADDR, depth_addr=DEPTH_ADDR, width=width, height=height, color_size=size,
                  num_triangles=n)


def batch_streams(count: int = 64, **kw) -> List[Stream]:
    return [batch_stream(i, **kw) for i in range(count)]


def split_at_sync_full(words: np.ndarray) -> List[np.ndarray]:
    """Splits a command stream into frames that each end with a SyncFull
    (trailing commands after the last SyncFull form a final frame)."""
    frames = []
    start = i = 0
    n = int(words.size)
    while i < n:
        op = (int(words[i]) >> 24) & 0x3F
        i += int(COMMAND_LENGTHS[op]) or 1
        if op == 0x29:
            frames.append(words[start:i])
            start = i
    if start < n:
        frames.append(words[start:n])
    return frames


# ----------------------------------------------------------------------------
# Random-state streams: every mode bit, selector, format and primitive kind
# drawn at random on a small framebuffer, for parity testing.
# ----------------------------------------------------------------------------
def _set_key_r(cl: CommandList, width: int, center: int, scale: int):
    cl.raw([0x2B << 24, (width & 0xFFF) << 16 | (center & 0xFF) << 8 | (scale & 0xFF)])


def _set_key_gb(cl: CommandList, width_g: int, width_b: int, center_g: int, scale_g: int, center_b: int,
                scale_b: int):
    cl.raw([0x2A << 24 | (width_g & 0xFFF) << 12 | (width_b & 0xFFF),
            (center_g & 0xFF) << 24 | (scale_g & 0xFF) << 16 | (center_b & 0xFF) << 8 | (scale_b & 0xFF)])


# (fmt, size) pairs a tile or texture image can take.
_TEX_FORMATS = [(FMT_RGBA, SIZ_16), (FMT_RGBA, SIZ_32), (FMT_YUV, SIZ_16), (FMT_CI, SIZ_4), (FMT_CI, SIZ_8),
                (FMT_IA, SIZ_4), (FMT_IA, SIZ_8), (FMT_IA, SIZ_16), (FMT_I, SIZ_4), (FMT_I, SIZ_8),
                (FMT_RGBA, SIZ_8), (FMT_I, SIZ_16)]
# (fmt, size) of colour images: I4, I8, RGBA5551, IA88, RGBA8888.
_FB_FORMATS = [(FMT_I, SIZ_4), (FMT_I, SIZ_8), (FMT_RGBA, SIZ_16), (FMT_IA, SIZ_16), (FMT_RGBA, SIZ_32)]


def _random_modes(rng, cycle_type: int) -> OtherModes:
    b = lambda p=0.5: bool(rng.random() < p)
    r4 = lambda: int(rng.integers(0, 4))
    return OtherModes(
        cycle_type=cycle_type, perspective=b(), detail_lod=b(0.2), sharpen_lod=b(0.2), tex_lod=b(0.3),
        tlut=b(0.3), tlut_ia=b(0.3), sample_quad=b(), mid_texel=b(0.2), bilerp0=b(), bilerp1=b(),
        convert_one=b(0.2), rgb_dither=r4(), alpha_dither=r4(),
        blend=tuple(tuple(r4() for _ in range(4)) for _ in range(2)), force_blend=b(), alpha_cvg_select=b(0.3),
        cvg_times_alpha=b(0.3), z_mode=r4(), coverage_mode=r4(), color_on_coverage=b(0.3), image_read=b(),
        depth_write=b(), depth_test=b(), aa=b(), prim_depth=b(0.2), alpha_test_dither=b(0.2), alpha_test=b(0.3))


def _random_combine(rng, cl: CommandList):
    sel = lambda n: int(rng.integers(0, n))
    c = [(sel(16), sel(16), sel(32), sel(8)) for _ in range(2)]
    a = [(sel(8), sel(8), sel(8), sel(8)) for _ in range(2)]
    cl.set_combine(c[0], a[0], c[1], a[1])


def _random_colors(rng, cl: CommandList):
    u32 = lambda: int(rng.integers(0, 1 << 32, dtype=np.int64))
    cl.set_prim_color(u32(), min_level=int(rng.integers(0, 32)), lod_frac=int(rng.integers(0, 256)))
    cl.set_env_color(u32())
    cl.set_fog_color(u32())
    cl.set_blend_color(u32())
    cl.set_fill_color(u32())
    cl.set_prim_depth(int(rng.integers(0, 1 << 16)), int(rng.integers(0, 1 << 16)))
    cl.set_convert([int(v) for v in rng.integers(0, 512, size=6)])
    _set_key_r(cl, int(rng.integers(0, 4096)), int(rng.integers(0, 256)), int(rng.integers(0, 256)))
    _set_key_gb(cl, *[int(rng.integers(0, 4096)) for _ in range(2)], *[int(rng.integers(0, 256)) for _ in range(4)])


def _random_tiles(rng, cl: CommandList):
    for tile in range(8):
        fmt, size = _TEX_FORMATS[int(rng.integers(0, len(_TEX_FORMATS)))]
        cl.set_tile(tile, fmt, size, int(rng.integers(0, 512)) * 8, int(rng.integers(0, 48)) * 8,
                    palette=int(rng.integers(0, 16)), mask_s=int(rng.integers(0, 9)), shift_s=int(rng.integers(0, 16)),
                    mask_t=int(rng.integers(0, 9)), shift_t=int(rng.integers(0, 16)), clamp_s=bool(rng.random() < 0.5),
                    mirror_s=bool(rng.random() < 0.3), clamp_t=bool(rng.random() < 0.5),
                    mirror_t=bool(rng.random() < 0.3))
        s0, t0 = (int(v) for v in rng.integers(0, 64, size=2))
        s1, t1 = s0 + int(rng.integers(0, 192)), t0 + int(rng.integers(0, 192))
        cl.set_tile_size(tile, s0, t0, s1, t1)


def _random_load(rng, cl: CommandList):
    fmt, size = _TEX_FORMATS[int(rng.integers(0, len(_TEX_FORMATS)))]
    width = int(rng.integers(1, 128))
    cl.set_texture_image(TEXTURE_ADDR + int(rng.integers(0, 4096)) * 8, width, size, fmt)
    tile = int(rng.integers(0, 8))
    tfmt, tsize = _TEX_FORMATS[int(rng.integers(0, len(_TEX_FORMATS)))]
    cl.set_tile(tile, tfmt, tsize, int(rng.integers(0, 512)) * 8, int(rng.integers(0, 40)) * 8,
                palette=int(rng.integers(0, 16)))
    kind = int(rng.integers(0, 3))
    s0, t0 = int(rng.integers(0, 64)), int(rng.integers(0, 64))
    if kind == 0:
        cl.load_tile(tile, s0, t0, s0 + int(rng.integers(0, 160)), t0 + int(rng.integers(0, 96)))
    elif kind == 1:
        cl.load_block(tile, s0 >> 2, t0 >> 2, (s0 >> 2) + int(rng.integers(0, 2048)), int(rng.integers(0, 4096)))
    else:
        cl.load_tlut(tile, s0, t0, s0 + (int(rng.integers(0, 256)) << 2), t0 + (4 if rng.random() < 0.1 else 0))
    cl.set_tile_size(tile, 0, 0, int(rng.integers(4, 256)), int(rng.integers(4, 256)))


def _random_triangles(rng, cl: CommandList, n: int, width: int, height: int):
    op_bits = int(rng.integers(0, 8))
    shade, tex, zbuf = bool(op_bits & 4), bool(op_bits & 2), bool(op_bits & 1)
    xy = _log_uniform_triangles(rng, n, width, height, 1.0, 1.5 * max(width, height))
    xy[:, :, 0] += rng.integers(-40, 40, size=(n, 1))
    rgba = rng.uniform(-40.0, 300.0, size=(n, 3, 4)) if shade else None
    z = rng.uniform(-2000.0, 0x40000 + 2000.0, size=(n, 3)) if zbuf else None
    uv = rng.uniform(-64.0, 160.0, size=(n, 3, 2)) if tex else None
    w = rng.uniform(0.05, 1.0, size=(n, 3)) if tex else None
    persp = bool(tex and rng.random() < 0.6)
    tris = triangle_commands(xy, rgba=rgba, z=z, uv=uv, w=w, perspective=persp,
                             tile=int(rng.integers(0, 8)), max_level=int(rng.integers(0, 8)))
    if tris.shape[0] and rng.random() < 0.3:
        # Perturb the edge slopes and starting points a little.
        k = tris.shape[0]
        for col in (2, 3, 4, 5, 6, 7):
            noise = rng.integers(-0x8000, 0x8000, size=k).astype(np.int64)
            tris[:, col] = ((tris[:, col].astype(np.int64) + noise) & 0xFFFFFFFF).astype(np.uint32)
    cl.extend_array(tris)
    return int(tris.shape[0])


def _random_rects(rng, cl: CommandList, n: int, width: int, height: int):
    for _ in range(n):
        x0, y0 = int(rng.integers(-16, width * 4)), int(rng.integers(-16, height * 4))
        x1, y1 = x0 + int(rng.integers(0, width * 4)), y0 + int(rng.integers(0, height * 3))
        x0, y0 = max(x0, 0), max(y0, 0)
        if rng.random() < 0.4:
            cl.fill_rectangle(x0, y0, x1, y1)
        else:
            cl.texture_rectangle(int(rng.integers(0, 8)), x0, y0, x1, y1, int(rng.integers(-2048, 8192)),
                                 int(rng.integers(-2048, 8192)), int(rng.integers(-4096, 4096)),
                                 int(rng.integers(-4096, 4096)), flip=bool(rng.random() < 0.3))


def fuzz_stream(seed: int, segments: int = 12, width: Optional[int] = None, height: Optional[int] = None) -> Stream:
    """A stream whose draw state is re-randomised every segment: cycle type,
    combiner and blender selectors, depth/coverage/dither modes, constant
    colours, tile descriptors, TMEM loads (tile, block, TLUT), triangles of
    every opcode and rectangles.  RDRAM starts random; the framebuffer format
    is one of I4, I8, RGBA5551, IA88, RGBA8888."""
    rng = np.random.default_rng(seed)
    width = width or int(rng.choice([32, 64, 96, 160]))
    height = height or int(rng.choice([24, 48, 72, 120]))
    fb_fmt, fb_size = _FB_FORMATS[int(rng.integers(0, len(_FB_FORMATS)))]
    rdram = rng.integers(0, 1 << 32, size=RDRAM_SIZE // 4, dtype=np.uint64).astype(np.uint32)
    cl = CommandList()
    cl.set_color_image(COLOR_ADDR, width, fb_size, fb_fmt)
    cl.set_mask_image(DEPTH_ADDR)
    total = 0
    for _ in range(segments):
        if rng.random() < 0.3:
            sx0, sy0 = int(rng.integers(0, width * 2)), int(rng.integers(0, height * 2))
            cl.set_scissor(sx0, sy0, sx0 + int(rng.integers(0, width * 4 + 64)),
                           sy0 + int(rng.integers(0, height * 4 + 64)), field=bool(rng.random() < 0.2),
                           keep_odd=bool(rng.random() < 0.5))
        else:
            cl.set_scissor(0, 0, width << 2, height << 2)
        cycle = int(rng.choice([0, 1, 2, 3], p=[0.4, 0.35, 0.1, 0.15]))
        cl.set_other_modes(_random_modes(rng, cycle))
        _random_combine(rng, cl)
        _random_colors(rng, cl)
        _random_tiles(rng, cl)
        for _ in range(int(rng.integers(0, 3))):
            _random_load(rng, cl)
        _random_tiles(rng, cl) if rng.random() < 0.3 else None
        if cycle >= 2:
            _random_rects(rng, cl, int(rng.integers(1, 6)), width, height)
            total += 1
        else:
            total += _random_triangles(rng, cl, int(rng.integers(1, 24)), width, height)
            if rng.random() < 0.5:
                _random_rects(rng, cl, int(rng.integers(1, 4)), width, height)
        if rng.random() < 0.3:
            cl.sync_full()
    cl.sync_full()
    return Stream(name=f"fuzz{seed:#x}:{width}x{height}:{fb_fmt}/{fb_size}", seed=seed, frames=[Frame(cl.words())],
                  rdram_init=rdram, hidden_init=int(rng.integers(0, 4)), color_addr=COLOR_ADDR,
                  depth_addr=DEPTH_ADDR, width=width, height=height, color_size=fb_size, num_triangles=total)


def alias_stream(seed: int, fb_fmt: int = FMT_RGBA, cycle: int = 0, partial: bool = False, segments: int = 8) -> Stream:
    """Colour and depth images that share memory, drawn with the depth test
    and update on (16-bit images: RGBA5551 or IA88).

    partial=False: the depth image is the colour image, the reference's
    colour/depth alias (memory_interfacing.h:100-151, 484-571).
    partial=True: the depth image starts half the image height into the
    colour image.  Each segment draws inside the top or the bottom half
    (scissor) and ends with a SyncFull, so within one render pass the colour
    and depth addresses written never meet (two tiles storing to one address
    in a pass would race in the reference too), while later passes read
    colour memory that holds depth values and the other way round."""
    rng = np.random.default_rng(seed)
    width = int(rng.choice([32, 64, 96]))
    height = int(rng.choice([32, 48, 64]))
    rdram = rng.integers(0, 1 << 32, size=RDRAM_SIZE // 4, dtype=np.uint64).astype(np.uint32)
    depth_addr = COLOR_ADDR + ((height // 2) * width * 2 if partial else 0)
    cl = CommandList()
    cl.set_color_image(COLOR_ADDR, width, SIZ_16, fb_fmt)
    cl.set_mask_image(depth_addr)
    total = 0
    for seg in range(segments):
        if partial:
            half = height // 2
            y0 = 0 if seg % 2 == 0 else half
            cl.set_scissor(0, y0 << 2, width << 2, (y0 + half) << 2)
        else:
            cl.set_scissor(0, 0, width << 2, height << 2)
        modes = _random_modes(rng, cycle)
        modes.depth_test = True
        modes.depth_write = bool(rng.random() < 0.8)
        cl.set_other_modes(modes)
        _random_combine(rng, cl)
        _random_colors(rng, cl)
        _random_tiles(rng, cl)
        if rng.random() < 0.5:
            _random_load(rng, cl)
        total += _random_triangles(rng, cl, int(rng.integers(4, 24)), width, height)
        if rng.random() < 0.3:
            _random_rects(rng, cl, int(rng.integers(1, 3)), width, height)
        if partial or rng.random() < 0.4:
            cl.sync_full()
    cl.sync_full()
    kind = "partial" if partial else "alias"
    return Stream(name=f"{kind}{seed:#x}:{width}x{height}:{fb_fmt}:c{cycle}", seed=seed, frames=[Frame(cl.words())],
                  rdram_init=rdram, hidden_init=int(rng.integers(0, 4)), color_addr=COLOR_ADDR, depth_addr=depth_addr,
                  width=width, height=height, color_size=SIZ_16, num_triangles=total)
