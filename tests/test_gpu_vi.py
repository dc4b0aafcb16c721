"""VI scanout on the device against the oracle: the reference's
per-scanline-xh known answer, random register/option sequences (every
filter, format, PAL/NTSC, serrate, crop, blending and persistence), and
scanout of rendered frames."""
import numpy as np
import pytest

import harness
import rdp_streams as rs
import test_oracle_vi as tv
import vi_cases

pytestmark = pytest.mark.gpu


def gpu_opts(d):
    import rdp_mi
    return rdp_mi.ScanoutOptions.default(**d)


def oracle_opts(d):
    import oracle
    return oracle.ScanoutOptions.default(**d)


@pytest.mark.parametrize("crop", [None, (9, 11, 10, 12)])
def test_per_scanline_xh_known_answer_gpu(crop):
    import rdp_mi
    p = rdp_mi.CommandProcessor(rs.RDRAM_SIZE)
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
    img = p.scanout_sync(gpu_opts(kw))
    ref = tv.per_scanline_reference(crop)
    assert img.shape == ref.shape
    assert np.array_equal((img & 0xFF).astype(np.int64), ref)
    assert not ((img >> 8) & 0xFFFF).any()
    p.close()


class _ScanoutAdapter:
    """Lets the oracle-side helper drive a device processor."""

    def __init__(self, p):
        self.p = p

    def set_vi_register(self, r, v):
        self.p.set_vi_register(r, v)

    def scanout(self, opts=None):
        import rdp_mi
        d = {f: getattr(opts, f) for f in vi_cases.OPTION_FIELDS} if opts is not None else {}
        return self.p.scanout_sync(rdp_mi.ScanoutOptions.default(**d))


@pytest.mark.parametrize("seed", range(0, 48, 6))
def test_random_vi_sequences_match_oracle(oracle_mod, seed):
    import rdp_mi
    for case in range(seed, seed + 6):
        rdram, hidden, frames = vi_cases.random_case(case)
        p = rdp_mi.CommandProcessor(rs.RDRAM_SIZE)
        p.write_rdram(rdram)
        p.write_hidden_rdram(hidden)
        o = oracle_mod.Oracle(rs.RDRAM_SIZE)
        o.rdram[:] = rdram
        o.hidden[:] = hidden
        for k, (regs, opts, per_line) in enumerate(frames):
            print(f"[case] vi seed {case} frame {k} control {regs[rs.VI_CONTROL]:#x} opts {opts}", flush=True)
            vi_cases.apply_frame(p, regs, per_line)
            vi_cases.apply_frame(o, regs, per_line)
            got = p.scanout_sync(gpu_opts(opts))
            exp = o.scanout(oracle_opts(opts))
            assert got.shape == exp.shape, (case, k, got.shape, exp.shape)
            diff = np.nonzero(got != exp)
            assert diff[0].size == 0, (case, k, diff[0].size, int(diff[0][0]), int(diff[1][0]),
                                       hex(int(got[diff][0])), hex(int(exp[diff][0])))
        o.close()
        p.close()


def test_scanout_of_rendered_frame(oracle_mod):
    """Render a config-0 frame, then scan it out with AA + divot + dither
    filter + gamma on both sides."""
    import rdp_mi
    s = rs.config0_stream()
    p = rdp_mi.CommandProcessor(rs.RDRAM_SIZE)
    gpu = harness.run_gpu(s, proc=p)
    o = oracle_mod.Oracle(rs.RDRAM_SIZE)
    o.hidden[:] = s.hidden_init
    for f in s.frames:
        o.process(f.words)
    regs = {rs.VI_CONTROL: 2 | (1 << 8) | (1 << 4) | (1 << 16) | (1 << 3), rs.VI_ORIGIN: rs.COLOR_ADDR,
            rs.VI_WIDTH: 320, rs.VI_V_SYNC: 525, rs.VI_V_START: vi_cases.start_reg(34 + 4, 34 + 484),
            rs.VI_H_START: vi_cases.start_reg(108, 108 + 640), rs.VI_X_SCALE: vi_cases.scale_reg(512, 0),
            rs.VI_Y_SCALE: vi_cases.scale_reg(1024, 0)}
    for r, v in regs.items():
        p.set_vi_register(r, v)
        o.set_vi_register(r, v)
    got = p.scanout_sync()
    exp = o.scanout()
    assert got.shape == (240, 640) and np.array_equal(got, exp)
    assert (got != 0).any()
    o.close()
    p.close()
