"""Closed-form and brute-force pins for individual functions of the oracle
(z_encode.h, clamping.h, combiner.h, coverage.h, noise.h, dither.h,
perspective.h, rdp_renderer.cpp:717-740)."""
import ctypes
import math

import numpy as np
import pytest


@pytest.fixture(scope="module")
def lib(oracle_mod):
    L = oracle_mod._load()
    L.rdpo_test_z_compress.restype = ctypes.c_uint32
    L.rdpo_test_z_compress.argtypes = [ctypes.c_int32]
    L.rdpo_test_z_decompress.restype = ctypes.c_int32
    L.rdpo_test_z_decompress.argtypes = [ctypes.c_uint32]
    for n in ("rdpo_test_clamp_9bit", "rdpo_test_clamp_z", "rdpo_test_dz_compress_host", "rdpo_test_normalize_dzpix"):
        getattr(L, n).restype = ctypes.c_int32
        getattr(L, n).argtypes = [ctypes.c_int32]
    L.rdpo_test_combiner_equation.restype = ctypes.c_int32
    L.rdpo_test_combiner_equation.argtypes = [ctypes.c_int32] * 4
    L.rdpo_test_noise.restype = ctypes.c_uint32
    L.rdpo_test_noise.argtypes = [ctypes.c_uint32] * 3
    L.rdpo_test_perspective_divide.argtypes = [ctypes.c_int32] * 3 + [ctypes.c_void_p]
    L.rdpo_test_coverage.restype = ctypes.c_int
    L.rdpo_test_coverage.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int]
    L.rdpo_test_rgb_dither.argtypes = [ctypes.c_void_p, ctypes.c_int]
    return L


def test_z_codec_exhaustive(lib):
    # Every 14-bit code decompresses into the 18-bit range, monotonically,
    # and compress(decompress(code)) == code.
    prev = -1
    for code in range(1 << 14):
        z = lib.rdpo_test_z_decompress(code)
        assert 0 <= z <= 0x3FFFF
        assert z > prev
        prev = z
        assert lib.rdpo_test_z_compress(z) == code
    # compress is monotone and decompress(compress(z)) <= z for all 2^18 z.
    last = 0
    for z in range(0, 1 << 18, 7):
        c = lib.rdpo_test_z_compress(z)
        assert c >= last
        last = c
        assert lib.rdpo_test_z_decompress(c) <= z
    # Exponent boundaries follow the leading ones of 0x3ffff - z.
    assert lib.rdpo_test_z_compress(0) == 0
    assert lib.rdpo_test_z_compress(0x3FFFF) == (7 << 11) | 0x7FF


def test_clamp_9bit_exhaustive(lib):
    for c in range(-1024, 1024):
        v = lib.rdpo_test_clamp_9bit(c)
        low9 = ((c - 0x80) & 0x1FF)
        s9 = low9 - 0x200 if low9 & 0x100 else low9
        assert v == min(max(s9 + 0x80, 0), 255)


def test_clamp_z_ranges(lib):
    assert lib.rdpo_test_clamp_z(0) == 0
    assert lib.rdpo_test_clamp_z(0x3FFFF) == 0x3FFFF
    for z in range(0x40000, 0x60000, 97):
        assert lib.rdpo_test_clamp_z(z) == 0x3FFFF
    for z in range(0x60000, 0x80000, 97):
        assert lib.rdpo_test_clamp_z(z) == 0


def test_dz_compress_and_normalize(lib):
    for k in range(16):
        assert lib.rdpo_test_dz_compress_host(1 << k) == k
    assert lib.rdpo_test_normalize_dzpix(0) == 1
    assert lib.rdpo_test_normalize_dzpix(0x8000) == 0x8000
    for dz in range(1, 0x8000, 13):
        v = lib.rdpo_test_normalize_dzpix(dz)
        assert v == 1 << (dz.bit_length())


def test_combiner_equation_grid(lib):
    def sx9(v):
        v &= 0x1FF
        return v - 0x200 if v & 0x100 else v

    def expand(v):
        return sx9(v - 0x80) + 0x80

    rng = np.random.default_rng(7)
    for _ in range(20000):
        a, b, c, d = (int(x) for x in rng.integers(-300, 600, size=4))
        ref = ((expand(a) - expand(b)) * sx9(c) + 0x80 >> 8) + expand(d)
        assert lib.rdpo_test_combiner_equation(a, b, c, d) == ref
    # (a - b) * 1.0 + 0 == a - b for in-range values with c == 0x100 -> sign bit: c is 9-bit signed
    assert lib.rdpo_test_combiner_equation(200, 0, 0xFF, 0) == (200 * 255 + 0x80) >> 8


def test_coverage_brute_force(lib):
    # Eight samples per pixel at x offsets 0,4 / 2,6 / 0,4 / 2,6 (eighths of a
    # pixel) on the four sub-scanlines; a sample is covered when it lies in
    # [xleft, xright) of its sub-scanline.
    rng = np.random.default_rng(3)
    offs = [0, 4, 2, 6, 0, 4, 2, 6]
    for _ in range(3000):
        xl = rng.integers(0, 200, size=4).astype(np.uint16)
        xr = (xl + rng.integers(0, 60, size=4)).astype(np.uint16)
        x = int(rng.integers(0, 30))
        got = lib.rdpo_test_coverage(xl.ctypes.data, xr.ctypes.data, x)
        exp = 0
        for k in range(8):
            line = k >> 1
            xs = x * 8 + offs[k]
            if xl[line] <= xs < xr[line]:
                exp |= 1 << k
        assert got == exp


def test_noise_hash_matches_shadertoy_form(lib):
    P = 1103515245
    for x, y, p in [(0, 0, 0), (1, 2, 3), (639, 479, 4095), (12345, 54321, 99)]:
        s = [x, y, p]
        for _ in range(3):
            s = [(((s[0] >> 8) ^ s[1]) * P) & 0xFFFFFFFF, (((s[1] >> 8) ^ s[2]) * P) & 0xFFFFFFFF,
                 (((s[2] >> 8) ^ s[0]) * P) & 0xFFFFFFFF]
        assert lib.rdpo_test_noise(x, y, p) == s[0] >> 16


def test_perspective_divide_tracks_exact_quotient(lib):
    # s16 divided by w in s1.15: the result approximates s * 2^15 / w to
    # within the reciprocal table's interpolation error; w <= 0 saturates.
    out = (ctypes.c_int32 * 3)()
    rng = np.random.default_rng(11)
    for w in list(range(0x400, 0x8000, 0x3F7)) + [0x4000, 0x2000, 0x7FFF]:
        for s in [int(v) for v in rng.integers(-2000, 2000, size=6)]:
            lib.rdpo_test_perspective_divide(s, s, w, out)
            exact = s * 32768 / w
            if abs(exact) >= 0x7FFF:
                continue
            assert abs(out[0] - exact) <= max(1.5, abs(exact) * 0.002), (s, w, out[0], exact)
            assert out[2] == 0
    # powers of two are exact
    lib.rdpo_test_perspective_divide(100, -100, 0x4000, out)
    assert out[0] == 200 and out[1] == -200
    lib.rdpo_test_perspective_divide(10, 10, 0, out)
    assert out[0] == 0x7FFF and out[2] == 1


def test_rgb_dither_closed_form(lib):
    arr = (ctypes.c_int32 * 3)()
    for c in range(256):
        for d in range(8):
            arr[0] = arr[1] = arr[2] = c
            lib.rdpo_test_rgb_dither(arr, d * 0x49)
            if (c & 7) > d:
                exp = 255 if c > 247 else (c & 0xF8) + 8
            else:
                exp = c
            assert arr[0] == exp & 0xFF, (c, d, arr[0])
