"""Pins the oracle's TMEM upload (a per-halfword gather restating
tmem_update.comp) against an independent forward model of the RDP load
pipeline: each 64-bit load iteration writes four 16-bit TMEM words, with the
32-bit word swap on odd lines (LoadTile) or odd T (LoadBlock), and LoadTLUT
replicating each palette entry into four consecutive words.  RDRAM holds the
ramp pattern of the reference's load conformance test (rdp_conformance.cpp:
628-637)."""
import numpy as np
import pytest

import rdp_streams as rs

TEX_BASE = 2 * 1024 * 1024


def ramp_rdram():
    rd = np.zeros(rs.RDRAM_SIZE // 4, dtype=np.uint32)
    i = np.arange(64 * 1024, dtype=np.uint32)
    v = ((4 * i & 0xFF) << 24) | (((4 * i + 1) & 0xFF) << 16) | (((4 * i + 2) & 0xFF) << 8) | ((4 * i + 3) & 0xFF)
    rd[512 * 1024: 512 * 1024 + 64 * 1024] = v
    return rd


def rd16(rdram, addr):
    h = rdram.view(np.uint16)
    return int(h[((addr >> 1) ^ 1)])


def run_load(oracle_mod, words):
    o = oracle_mod.Oracle(rs.RDRAM_SIZE)
    o.rdram[:] = ramp_rdram()
    o.process(np.asarray(words, dtype=np.uint32))
    o.flush()
    t = o.tmem.copy()
    rdram = o.rdram.copy()
    o.close()
    return t, rdram


def logical(tmem):
    """TMEM as logical halfwords (undo the ^1 host swizzle)."""
    return tmem[np.arange(2048) ^ 1]


def cmds_tile(width, height, offset, stride, size, tex_width, rdram_offset=0):
    cl = rs.CommandList()
    cl.set_texture_image(TEX_BASE + rdram_offset, tex_width, size, rs.FMT_RGBA)
    cl.set_tile(0, rs.FMT_RGBA, size, offset, stride)
    cl.load_tile(0, 0, 0, (width - 1) << 2, (height - 1) << 2)
    return cl.words()


@pytest.mark.parametrize("size", [rs.SIZ_8, rs.SIZ_16])
@pytest.mark.parametrize("width,height,offset,stride", [
    (32, 7, 0, 128), (32, 7, 0x800, 128), (32, 7, 128, 128), (16, 3, 0, 256), (12, 5, 64, 64),
    (8, 3, 0, 8), (20, 2, 0, 48), (64, 8, 0, 136)])
def test_load_tile_forward_model(oracle_mod, size, width, height, offset, stride):
    tex_width = 64
    tmem, rdram = run_load(oracle_mod, cmds_tile(width, height, offset, stride, size, tex_width))
    exp = np.zeros(2048, dtype=np.uint16)
    bpp = 1 << (size - 1)
    words_per_line = ((width * bpp + 7) & ~7) // 2
    for y in range(height):
        line = TEX_BASE + y * tex_width * bpp
        for x in range(words_per_line):
            pos = ((offset >> 1) + y * (stride >> 1) + (x ^ ((y & 1) << 1))) & 0x7FF
            exp[pos] = rd16(rdram, line + 2 * x)
    assert np.array_equal(logical(tmem), exp)


@pytest.mark.parametrize("pixels,dxt,offset", [(64, 0, 0), (256, 0x800, 0), (512, 0x400, 0x100), (2048, 0x200, 0),
                                               (100, 0x155, 0), (1024, 0x100, 0x800)])
def test_load_block_forward_model(oracle_mod, pixels, dxt, offset):
    cl = rs.CommandList()
    cl.set_texture_image(TEX_BASE, 256, rs.SIZ_16, rs.FMT_RGBA)
    cl.set_tile(0, rs.FMT_RGBA, rs.SIZ_16, offset, 0)
    cl.load_block(0, 0, 0, pixels - 1, dxt)
    tmem, rdram = run_load(oracle_mod, cl.words())
    exp = np.zeros(2048, dtype=np.uint16)
    n_words = (pixels + 3) & ~3
    for q in range(n_words):
        t = ((q >> 2) * dxt) >> 11
        pos = ((offset >> 1) + (q ^ ((t & 1) << 1))) & 0x7FF
        exp[pos] = rd16(rdram, TEX_BASE + 2 * q)
    assert np.array_equal(logical(tmem), exp)


@pytest.mark.parametrize("entries,offset", [(16, 0x800), (256, 0x800), (16, 0xC00), (7, 0x800)])
def test_load_tlut_forward_model(oracle_mod, entries, offset):
    cl = rs.CommandList()
    cl.set_texture_image(TEX_BASE, 256, rs.SIZ_16, rs.FMT_RGBA)
    cl.set_tile(7, rs.FMT_RGBA, rs.SIZ_4, offset, 0)
    cl.load_tlut(7, 0, 0, (entries - 1) << 2, 0)
    tmem, rdram = run_load(oracle_mod, cl.words())
    exp = np.zeros(2048, dtype=np.uint16)
    for e in range(entries):
        for k in range(4):
            exp[((offset >> 1) + 4 * e + k) & 0x7FF] = rd16(rdram, TEX_BASE + 2 * e)
    assert np.array_equal(logical(tmem), exp)


def test_load_tile_rgba32_splits_halves(oracle_mod):
    # 32-bit texels: red/green halfword in low TMEM, blue/alpha in high TMEM.
    width, height, stride = 16, 4, 64
    cl = rs.CommandList()
    cl.set_texture_image(TEX_BASE, 64, rs.SIZ_32, rs.FMT_RGBA)
    cl.set_tile(0, rs.FMT_RGBA, rs.SIZ_32, 0, stride)
    cl.load_tile(0, 0, 0, (width - 1) << 2, (height - 1) << 2)
    tmem, rdram = run_load(oracle_mod, cl.words())
    lg = logical(tmem)
    exp = np.zeros(2048, dtype=np.uint16)
    for y in range(height):
        for x in range(width):
            addr = TEX_BASE + (y * 64 + x) * 4
            hi, lo = rd16(rdram, addr), rd16(rdram, addr + 2)
            pos = (y * (stride >> 1) + (x ^ ((y & 1) << 1))) & 0x3FF
            exp[pos] = hi
            exp[pos | 0x400] = lo
    assert np.array_equal(lg, exp)
