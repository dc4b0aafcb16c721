"""Capture inspection tool on the CPU: its PNG encoder round-trips through an
independent PNG reader, and it follows SetColorImage / SetMaskImage the way
the reference's replayer does (rdp_replayer.cpp:395-421)."""
import struct
import zlib

import numpy as np

import rdp_replay
import rdp_streams as rs


def read_png(data: bytes) -> np.ndarray:
    assert data[:8] == b"\x89PNG\r\n\x1a\n"
    pos, idat, w = 8, b"", None
    while pos < len(data):
        n = struct.unpack(">I", data[pos:pos + 4])[0]
        tag, body = data[pos + 4:pos + 8], data[pos + 8:pos + 8 + n]
        assert struct.unpack(">I", data[pos + 8 + n:pos + 12 + n])[0] == zlib.crc32(tag + body) & 0xFFFFFFFF
        if tag == b"IHDR":
            w, h, depth, ctype = struct.unpack(">IIBB", body[:10])
            assert depth == 8 and ctype == 6
        elif tag == b"IDAT":
            idat += body
        pos += 12 + n
    raw = np.frombuffer(zlib.decompress(idat), np.uint8).reshape(h, 1 + 4 * w)
    assert not raw[:, 0].any()  # filter type 0 on every row
    return raw[:, 1:].copy().view("<u4").reshape(h, w)


def test_png_round_trip():
    rng = np.random.default_rng(3)
    img = rng.integers(0, 1 << 32, size=(37, 53), dtype=np.uint64).astype(np.uint32)
    assert np.array_equal(read_png(rdp_replay.png_bytes(img)), img)


def test_tracker_follows_color_and_mask_images():
    cl = rs.CommandList()
    cl.set_color_image(0x123456, 320, rs.SIZ_16, rs.FMT_RGBA)
    cl.set_mask_image(0x2A0000)
    cl.set_color_image(0x0ABC00, 640, rs.SIZ_32, rs.FMT_RGBA)
    words = cl.words()
    t = rdp_replay.FramebufferTracker()
    seen = []
    for i in range(0, words.size, 2):
        t.command((int(words[i]) >> 24) & 63, words[i:i + 2])
        seen.append((t.fb_addr, t.fb_size, t.fb_width, t.depth_addr))
    assert seen == [(0x123456, 2, 320, 0), (0x123456, 2, 320, 0x2A0000), (0x0ABC00, 3, 640, 0x2A0000)]
