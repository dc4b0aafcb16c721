"""Upscaled rendering on the GPU against the oracle: native RDRAM, hidden
RDRAM and TMEM, and every sample plane of RDRAM and hidden RDRAM, bit-exact
at 2x, 4x and 8x (RDP_MI_FLAG_UPSCALING_*); the VI at scale (the reference's
upscaled per-scanline known answer and scanouts of rendered frames); and
upscaled passes in batched recordings and graph replays."""
import numpy as np
import pytest

import rdp_streams as rs
import test_oracle_vi as tv

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module")
def rdp_mi():
    import rdp_mi as m
    m.load_library()
    return m


def _mismatch(a, b):
    a8, b8 = np.ascontiguousarray(a).view(np.uint8).ravel(), np.ascontiguousarray(b).view(np.uint8).ravel()
    d = np.nonzero(a8 != b8)[0]
    return int(d.size), (hex(int(d[0])) if d.size else "-")


NAMES = ("rdram", "hidden", "tmem", "upscaled rdram", "upscaled hidden")


def assert_same(gpu, ref, what):
    bad = []
    for name, g, r in zip(NAMES, gpu, ref):
        n, first = _mismatch(g, r)
        if n:
            bad.append(f"{name}: {n} bytes differ, first at byte {first}")
    assert not bad, f"{what}: " + "; ".join(bad)


def oracle_state(stream, factor, scanouts=None, opts=None, ssaa=None):
    """ssaa: None, or the dither switch of super-sampled readback."""
    import oracle
    o = oracle.Oracle(rs.RDRAM_SIZE)
    o.enable_upscaling(factor)
    if ssaa is not None:
        o.enable_super_sampled_readback(ssaa)
    if stream.rdram_init is not None:
        o.rdram[:] = stream.rdram_init
    o.hidden[:] = stream.hidden_init
    for f in stream.frames:
        o.process(f.words)
        if scanouts is not None and f.vi:
            for r, v in f.vi.items():
                o.set_vi_register(r, v)
            scanouts.append(o.scanout(opts))
    o.flush()
    out = (o.rdram.copy(), o.hidden.copy(), o.tmem.copy(), o.upscaled_rdram.copy(), o.upscaled_hidden.copy())
    o.close()
    return out


def new_proc(rdp_mi, stream, factor, hold=False, ssaa=None):
    flags = rdp_mi.FLAG_UPSCALING[factor]
    if ssaa is not None:
        flags |= rdp_mi.FLAG_SUPER_SAMPLED_READ_BACK | (rdp_mi.FLAG_SUPER_SAMPLED_DITHER if ssaa else 0)
    p = rdp_mi.CommandProcessor(rs.RDRAM_SIZE, flags=flags)
    assert p.upscaling_factor == factor
    if hold:
        p.hold_submissions()
    p.write_rdram(stream.rdram_init if stream.rdram_init is not None else np.zeros(rs.RDRAM_SIZE // 4, np.uint32))
    p.write_hidden_rdram(np.full(rs.RDRAM_SIZE // 2, stream.hidden_init, np.uint8))
    return p


def gpu_state(p):
    return (p.read_rdram(), p.read_hidden_rdram(), p.read_tmem(), p.read_upscaled_rdram(),
            p.read_upscaled_hidden_rdram())


def run_gpu(rdp_mi, stream, factor, scanouts=None, opts=None, ssaa=None):
    p = new_proc(rdp_mi, stream, factor, ssaa=ssaa)
    for f in stream.frames:
        p.enqueue_commands(f.words)
        if scanouts is not None and f.vi:
            for r, v in f.vi.items():
                p.set_vi_register(r, v)
            scanouts.append(p.scanout_sync(opts))
    p.idle()
    out = gpu_state(p)
    p.close()
    return out


@pytest.mark.parametrize("factor", [2, 4, 8])
def test_config0_upscaled(rdp_mi, factor):
    s = rs.config0_stream(num_triangles=64)
    say(f"[case] {s.name} x{factor}")
    assert_same(run_gpu(rdp_mi, s, factor), oracle_state(s, factor), f"{s.name} x{factor}")


@pytest.mark.parametrize("index,factor", [(0, 2), (1, 2), (2, 2), (3, 4), (5, 2), (6, 4), (9, 2), (12, 2)])
def test_batch_streams_upscaled(rdp_mi, index, factor):
    """Plain shaded, textured (TMEM snapshots shared with the native pass)
    and two-cycle CI-texture workloads, 320x240 and 640x480, 16 and 32 bit."""
    s = rs.batch_stream(index, max_tris=400)
    say(f"[case] {s.name} x{factor}")
    assert_same(run_gpu(rdp_mi, s, factor), oracle_state(s, factor), f"{s.name} x{factor}")


@pytest.mark.parametrize("seed", range(0, 24, 4))
def test_fuzz_streams_upscaled(rdp_mi, seed):
    """Random state per segment: every cycle type, combiner, blender, depth
    and coverage mode, TMEM loads, scissors with interlace, all five
    framebuffer formats (the native-only ones take only the pre update)."""
    for k in range(seed, seed + 4):
        s = rs.fuzz_stream(7000 + k, width=int((320, 288, 256, 400)[k % 4]), height=96)
        say(f"[case] {s.name} x2")
        assert_same(run_gpu(rdp_mi, s, 2), oracle_state(s, 2), f"{s.name} x2")


def test_native_resolution_quirk_off(rdp_mi):
    """Texture rectangles with the native-resolution quirk off render at
    scale (MetaSetQuirks, rdp_device.cpp:948-952)."""
    s = rs.batch_stream(4, max_tris=300)
    q = rs.CommandList()
    q.set_quirks(native_resolution_tex_rect=False, native_texture_lod=True)
    s.frames[0] = rs.Frame(np.concatenate([q.words(), s.frames[0].words]))
    say(f"[case] {s.name} quirks off x2")
    assert_same(run_gpu(rdp_mi, s, 2), oracle_state(s, 2), f"{s.name} quirks x2")


@pytest.mark.parametrize("crop", [None, (9, 11, 10, 12)])
def test_upscaled_per_scanline_known_answer(rdp_mi, crop):
    """run_per_scanline_xh_vi with upscale = true (vi_conformance.cpp:163-267):
    the frame is written to native RDRAM only and reaches the sample planes
    through the scanout's domain update."""
    p = rdp_mi.CommandProcessor(rs.RDRAM_SIZE, flags=rdp_mi.FLAG_UPSCALING[2])
    tv.default_vi(_ScanoutAdapter(p), fmt=tv.RGBA8888, aa=tv.AA_RESAMP_ONLY)
    fb = (np.tile(np.arange(200, dtype=np.uint32), 240) << 24).astype(np.uint32)
    rd = p.read_rdram()
    rd[4096 // 4: 4096 // 4 + fb.size] = fb
    p.write_rdram(rd)
    p.set_vi_register(rs.VI_ORIGIN, 4096)
    p.set_vi_register(rs.VI_WIDTH, 200)
    p.set_vi_register(rs.VI_V_START, tv.start_reg(tv.V_OFF + 40, tv.V_OFF + 400))
    p.begin_vi_register_per_scanline(3)
    for line, hs, xs in ((tv.V_OFF, tv.start_reg(tv.H_OFF, tv.H_OFF + 320), tv.scale_reg(256, 0)),
                         (tv.V_OFF + 100, tv.start_reg(tv.H_OFF, tv.H_OFF + 640), tv.scale_reg(240, 200)),
                         (tv.V_OFF + 200, tv.start_reg(tv.H_OFF - 8, tv.H_OFF + 648), tv.scale_reg(220, 400)),
                         (tv.V_OFF + 300, tv.start_reg(tv.H_OFF + 8, tv.H_OFF + 648), tv.scale_reg(210, 600))):
        p.set_vi_register_for_scanline(1, hs)
        p.set_vi_register_for_scanline(2, xs)
        p.latch_vi_register_for_scanline(line)
    p.end_vi_register_per_scanline()
    kw = dict(blend_previous_frame=1, upscale_deinterlacing=0)
    if crop:
        kw.update(crop_enable=1, crop_left=crop[0], crop_right=crop[1], crop_top=crop[2], crop_bottom=crop[3])
    img = p.scanout_sync(rdp_mi.ScanoutOptions.default(**kw))
    ref = tv.per_scanline_reference(crop, 2)
    assert img.shape == ref.shape
    assert np.array_equal((img & 0xFF).astype(np.int64), ref)
    assert not ((img >> 8) & 0xFFFF).any()
    p.close()


class _ScanoutAdapter:
    """Lets the oracle-side VI helper drive a device processor."""

    def __init__(self, p):
        self.p = p

    def set_vi_register(self, r, v):
        self.p.set_vi_register(r, v)

    def scanout(self, opts=None):
        import rdp_mi
        return self.p.scanout_sync(rdp_mi.ScanoutOptions.default(blend_previous_frame=1, upscale_deinterlacing=0))


@pytest.mark.parametrize("factor", [2, 4])
def test_upscaled_scanouts_of_rendered_frames(rdp_mi, factor):
    """Two-cycle frames with VI registers, scanned out at scale after each
    frame, with the reference conformance driver's options."""
    import oracle
    import rdp_mi as m
    s = rs.config2_stream(frames=2, prims_per_frame=300)
    kw = dict(blend_previous_frame=1, upscale_deinterlacing=0)
    say(f"[case] {s.name} scanouts x{factor}")
    g_scan, o_scan = [], []
    gpu = run_gpu(rdp_mi, s, factor, scanouts=g_scan, opts=m.ScanoutOptions.default(**kw))
    ref = oracle_state(s, factor, scanouts=o_scan, opts=oracle.ScanoutOptions.default(**kw))
    assert_same(gpu, ref, f"{s.name} x{factor}")
    assert len(g_scan) == len(o_scan) > 0
    for i, (a, b) in enumerate(zip(g_scan, o_scan)):
        assert a.shape == b.shape and a.shape[1] == 640 * factor, (i, a.shape, b.shape)
        assert np.array_equal(a, b), f"scanout {i}: {int(np.count_nonzero(a != b))} pixels differ"


def test_upscaled_graph_replays(rdp_mi):
    """Frames that start with a clear give the same native and upscaled
    state on every replay: one eager replay, then graph replays, each
    checked against the oracle's single run (2x, 4x and a readback
    processor in one batched recording)."""
    streams = [rs.batch_stream(i, max_tris=300) for i in (1, 4, 8)]
    modes = [(2, None), (4, None), (2, True)]
    say("[case] upscaled graph replays")
    procs = []
    for s, (f, ssaa) in zip(streams, modes):
        p = new_proc(rdp_mi, s, f, hold=True, ssaa=ssaa)
        for fr in s.frames:
            p.enqueue_commands(fr.words)
        procs.append(p)
    rec = rdp_mi.Recording(procs, batched=True)
    refs = [oracle_state(s, f, ssaa=ssaa) for s, (f, ssaa) in zip(streams, modes)]
    rec.replay(1, use_graph=False)
    for k in range(3):
        for p, s, ref in zip(procs, streams, refs):
            assert_same(gpu_state(p), ref, f"{s.name} replay {k}")
        rec.replay(1, use_graph=True)
    rec.close()
    for p in procs:
        p.close()


def test_upscaled_wide_rounds(rdp_mi):
    """Rounds of many passes run setup ahead on a second stream and the
    heavier raster classes beside the light one on a third; the domain's
    pre update must still see native RDRAM before any of the round's
    raster writes it (12 processors, mixed occupancy classes, 2x)."""
    streams = [rs.batch_stream(i, max_tris=300) for i in range(12)]
    say("[case] upscaled wide rounds")
    procs = []
    for s in streams:
        p = new_proc(rdp_mi, s, 2, hold=True)
        for fr in s.frames:
            p.enqueue_commands(fr.words)
        procs.append(p)
    rec = rdp_mi.Recording(procs, batched=True)
    rec.replay(1, use_graph=True)
    for p, s in zip(procs, streams):
        assert_same(gpu_state(p), oracle_state(s, 2), f"{s.name} x2")
    rec.close()
    for p in procs:
        p.close()


@pytest.mark.parametrize("batched", [True, False])
def test_upscaled_recording(rdp_mi, batched):
    """Upscaled passes follow their native passes round by round in a
    recording of several processors (some 2x, some 4x)."""
    streams = [rs.batch_stream(i, max_tris=300) for i in (0, 2, 7, 10)] + \
              [rs.fuzz_stream(7100 + i, width=320, height=64) for i in range(2)]
    factors = [2, 4, 2, 2, 4, 2]
    say(f"[case] upscaled recording batched={batched}")
    procs = []
    for s, f in zip(streams, factors):
        p = new_proc(rdp_mi, s, f, hold=True)
        for fr in s.frames:
            p.enqueue_commands(fr.words)
        procs.append(p)
    rec = rdp_mi.Recording(procs, batched=batched)
    rec.replay(1, use_graph=False)
    for p, s, f in zip(procs, streams, factors):
        assert_same(gpu_state(p), oracle_state(s, f), f"{s.name} x{f}")
    rec.close()
    for p in procs:
        p.close()


@pytest.mark.parametrize("factor,dither", [(2, False), (2, True), (4, False), (8, True)])
def test_super_sampled_readback_config0(rdp_mi, factor, dither):
    """RDP_MI_FLAG_SUPER_SAMPLED_READ_BACK: passes render at scale only and
    are resolved into native RDRAM."""
    s = rs.config0_stream(num_triangles=64)
    say(f"[case] {s.name} ssaa x{factor} dither={dither}")
    assert_same(run_gpu(rdp_mi, s, factor, ssaa=dither), oracle_state(s, factor, ssaa=dither),
                f"{s.name} ssaa x{factor}")


@pytest.mark.parametrize("index,factor,dither", [(1, 2, True), (2, 2, False), (3, 4, True), (6, 2, False)])
def test_super_sampled_readback_batch(rdp_mi, index, factor, dither):
    s = rs.batch_stream(index, max_tris=300)
    say(f"[case] {s.name} ssaa x{factor} dither={dither}")
    assert_same(run_gpu(rdp_mi, s, factor, ssaa=dither), oracle_state(s, factor, ssaa=dither),
                f"{s.name} ssaa x{factor}")


@pytest.mark.parametrize("seed", range(0, 12, 4))
def test_super_sampled_readback_fuzz(rdp_mi, seed):
    """Every framebuffer format and width renders at scale and resolves
    (8-bit colour from the first sample, depth from the first sample)."""
    for k in range(seed, seed + 4):
        s = rs.fuzz_stream(7300 + k, width=int((320, 160, 256, 96)[k % 4]), height=64)
        say(f"[case] {s.name} ssaa x2")
        assert_same(run_gpu(rdp_mi, s, 2, ssaa=bool(k & 1)), oracle_state(s, 2, ssaa=bool(k & 1)), f"{s.name} ssaa")


def test_super_sampled_readback_flags(rdp_mi):
    """Without an upscaling flag the readback flag is refused; with several
    upscaling flags the largest factor wins (rdp_device.cpp:188-195)."""
    with pytest.raises(rdp_mi.RDPError):
        rdp_mi.CommandProcessor(rs.RDRAM_SIZE, flags=rdp_mi.FLAG_SUPER_SAMPLED_READ_BACK)
    p = rdp_mi.CommandProcessor(rs.RDRAM_SIZE, flags=rdp_mi.FLAG_UPSCALING[2] | rdp_mi.FLAG_UPSCALING[4])
    assert p.upscaling_factor == 4
    p.close()


def fill_only_stream(seed, frames=3):
    """Passes of FILL-mode rectangles only (the upscaled fill kernel's work):
    random scissors, fractional rectangle edges, 8/16/32-bit colour images,
    IA88 fills and a colour image that is also the depth image."""
    rng = np.random.default_rng(seed)
    out = []
    width, height = 320, 240
    for _ in range(frames):
        cl = rs.CommandList()
        cl.set_other_modes(rs.OtherModes(cycle_type=3))
        for k in range(4):
            size, fmt = [(rs.SIZ_16, rs.FMT_RGBA), (rs.SIZ_32, rs.FMT_RGBA), (rs.SIZ_8, rs.FMT_I),
                         (rs.SIZ_16, rs.FMT_IA)][int(rng.integers(4))]
            addr = rs.DEPTH_ADDR if k == 3 else rs.COLOR_ADDR
            cl.set_color_image(addr, width, size, fmt)
            if k == 3:
                cl.set_mask_image(addr)
            x0, y0 = int(rng.integers(0, 40)), int(rng.integers(0, 40))
            cl.set_scissor(x0 << 2, y0 << 2, int(rng.integers(200, width + 1)) << 2,
                           int(rng.integers(150, height + 1)) << 2)
            for _ in range(int(rng.integers(1, 6))):
                cl.set_fill_color(int(rng.integers(0, 1 << 32)))
                a = rng.integers(0, (width << 2), 2)
                b = rng.integers(0, (height << 2), 2)
                cl.fill_rectangle(int(a.min()), int(b.min()), int(a.max()), int(b.max()))
            cl.sync_pipe()
        cl.sync_full()
        out.append(rs.Frame(words=cl.words()))
    return rs.Stream(name=f"fill_only_{seed}", seed=seed, frames=out, color_addr=rs.COLOR_ADDR,
                     depth_addr=rs.DEPTH_ADDR, width=width, height=height)


@pytest.mark.parametrize("seed,factor,ssaa", [(1, 2, None), (2, 4, None), (3, 8, None), (4, 2, True), (5, 4, False)])
def test_fill_only_passes_upscaled(rdp_mi, seed, factor, ssaa):
    stream = fill_only_stream(seed)
    say(f"fill-only stream {seed} at {factor}x, super-sampled readback {ssaa}")
    assert_same(run_gpu(rdp_mi, stream, factor, ssaa=ssaa), oracle_state(stream, factor, ssaa=ssaa),
                f"fill-only stream {seed} at {factor}x")
