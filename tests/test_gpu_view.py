"""Framebuffer views (rdp_mi_framebuffer_view) against the decoding the
reference's capture replayer applies (rdp_replayer.cpp:246-393), restated
here with numpy on RDRAM read back from the same processor; and the capture
replay tool end to end on a committed capture: its scanout images equal the
oracle's scanouts of the same capture."""
import os

import numpy as np
import pytest

import dumps
import rdp_streams as rs

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def rdp_mi():
    import rdp_mi as mod
    mod.load_library()
    return mod


def _expand5(c):
    return (c << 3) | (c >> 2)


def _rgba(r, g, b, a):
    return (r | (g << 8) | (b << 16) | (a << 24)).astype(np.uint32)


def _ref_views(rdram_words, hidden, fb_addr, fb_size, depth_addr, w, h):
    rd8 = rdram_words.view(np.uint8)
    rd16 = rdram_words.view(np.uint16)
    n = rdram_words.size * 4
    i = np.arange(w * h, dtype=np.int64)
    if fb_size == 2:
        idx = ((fb_addr >> 1) + i) & (n // 2 - 1)
        c = rd16[idx ^ 1].astype(np.uint32)
        color = _rgba(_expand5((c >> 11) & 31), _expand5((c >> 6) & 31), _expand5((c >> 1) & 31), (c & 1) * 0xff)
        cov = ((c & 1) << 2) | (hidden[idx] & 3)
        cov = ((cov << 5) | (cov << 2) | (cov >> 1)) & 0xff
        coverage = _rgba(cov, cov, cov, np.full_like(cov, 0xff))
    else:
        c = rdram_words[((fb_addr >> 2) + i) & (n // 4 - 1)].astype(np.uint32)
        color = _rgba(c >> 24, (c >> 16) & 0xff, (c >> 8) & 0xff, c & 0xff)
        coverage = np.zeros(w * h, np.uint32)
    idx = ((depth_addr >> 1) + i) & (n // 2 - 1)
    d = rd16[idx ^ 1].astype(np.uint32)
    dz = (hidden[idx].astype(np.uint32) | ((d & 3) << 2)) & 0xff
    d &= ~np.uint32(3)
    depth = _rgba(_expand5((d >> 11) & 31), _expand5((d >> 6) & 31), _expand5((d >> 1) & 31), dz)
    return [v.reshape(h, w) for v in (color, depth, coverage)]


@pytest.mark.parametrize("which", ["c0_16bit", "batch_32bit"])
def test_views_match_the_reference_decoding(rdp_mi, which):
    import harness
    stream = rs.config0_stream(num_triangles=64) if which == "c0_16bit" else rs.batch_stream(2, max_tris=300)
    print(f"[case] views of {stream.name}", flush=True)
    p = rdp_mi.CommandProcessor(rs.RDRAM_SIZE)
    harness.run_gpu(stream, proc=p)
    rd, hid = p.read_rdram(), p.read_hidden_rdram()
    fb_size = 2 if stream.color_size == rs.SIZ_16 else 3
    w, h = stream.width, stream.height
    ref = _ref_views(rd, hid, stream.color_addr, fb_size, stream.depth_addr, w, h)
    for mode, r in zip((p.VIEW_COLOR, p.VIEW_DEPTH, p.VIEW_COVERAGE), ref):
        got = p.framebuffer_view(mode, stream.color_addr, fb_size, stream.depth_addr, w, h)
        assert np.array_equal(got, r), f"view {mode}"
    # height 0: the reference's 4:3 assumption.
    assert p.framebuffer_view(p.VIEW_COLOR, stream.color_addr, fb_size, stream.depth_addr, w).shape == (w * 3 // 4, w)
    p.close()


def test_replay_tool_scanouts_match_the_oracle(rdp_mi, tmp_path):
    import oracle
    import rdp_replay
    from test_replay_tool import read_png
    path = dumps.golden_path("c2_two_cycle_320x240_vi")
    print("[case] replay tool on the two-cycle capture", flush=True)
    try:
        res = rdp_replay.replay(path, str(tmp_path), every_draws=100)
        assert res["frames"] == 2 and res["images"] > 2
        # The oracle replays the same capture and scans out at each EndFrame.
        rdram_size, _, recs = dumps.read_records(open(path, "rb").read())
    finally:
        os.remove(path)
    rep = dumps.OracleReplay(rdram_size)
    shadow = [np.zeros(rdram_size, np.uint8), np.zeros(rdram_size // 2, np.uint8)]
    opts = oracle.ScanoutOptions.default(blend_previous_frame=1, upscale_deinterlacing=0)
    frame = 0
    for kind, payload in recs:
        if kind in (dumps.UPDATE_DRAM, dumps.UPDATE_HIDDEN_DRAM):
            off, blk = payload
            shadow[kind == dumps.UPDATE_HIDDEN_DRAM][off:off + len(blk)] = np.frombuffer(blk, np.uint8)
        rep.apply(kind, payload, shadow[0].view(np.uint32), shadow[1])
        if kind == dumps.END_FRAME:
            ref = rep.o.scanout(opts)
            got = read_png(open(os.path.join(str(tmp_path), f"frame_{frame}_scanout.png"), "rb").read())
            assert np.array_equal(got, ref), f"frame {frame}"
            for v in ("color", "depth", "coverage"):
                assert os.path.exists(os.path.join(str(tmp_path), f"frame_{frame}_{v}.png"))
            frame += 1
    rep.close()
    assert frame == 2
