"""Passes of FILL-mode primitives only take their own raster kernel, which
reads nothing from the framebuffer and runs four tiles per wave.  These
streams render such passes in every colour-image format over random RDRAM
contents (so a pixel the fills miss must keep its old bytes), at widths that
are not a multiple of four tiles, under scissors, with enough rectangles for
two rounds of bin groups, and with fill passes next to shaded passes of the
same stream; RDRAM, hidden RDRAM and TMEM must match the oracle byte for
byte.  A batched recording of several such streams renders the fill passes
of all of them in one launch."""
import numpy as np
import pytest

import harness
import rdp_streams as rs

pytestmark = pytest.mark.gpu

FORMATS = [(rs.SIZ_16, rs.FMT_RGBA), (rs.SIZ_32, rs.FMT_RGBA), (rs.SIZ_8, rs.FMT_I), (rs.SIZ_16, rs.FMT_IA),
           (rs.SIZ_4, rs.FMT_I)]


def fill_stream(seed: int, many: bool = False) -> rs.Stream:
    rng = np.random.default_rng(seed)
    cl = rs.CommandList()
    cl.set_other_modes(rs.OtherModes(cycle_type=3))
    for k, (size, fmt) in enumerate(FORMATS):
        width = int(rng.choice([72, 200, 296, 320]))
        height = int(rng.integers(16, 120))
        cl.set_color_image(0x100000 + k * 0x100000, width, size, fmt)
        x0, y0 = int(rng.integers(0, 24)), int(rng.integers(0, 12))
        cl.set_scissor(x0 << 2, y0 << 2, (width - int(rng.integers(0, 24))) << 2, (height - int(rng.integers(0, 8))) << 2)
        n = 4200 if many and k == 0 else int(rng.choice([1, 5, 70, 600]))
        for _ in range(n):
            cl.set_fill_color(int(rng.integers(0, 1 << 32, dtype=np.int64)))
            fx = int(rng.integers(-8, width * 4))
            fy = int(rng.integers(-8, height * 4))
            w = int(rng.integers(0, (width * 4) // (8 if n > 100 else 2)))
            h = int(rng.integers(0, (height * 4) // (8 if n > 100 else 2)))
            cl.fill_rectangle(max(fx, 0), max(fy, 0), min(fx + w, 0xFFF), min(fy + h, 0xFFF))
        cl.sync_pipe()
    # A shaded pass, then fill passes again on the first image.
    cl.set_color_image(0x100000, 320, rs.SIZ_16, rs.FMT_RGBA)
    cl.set_scissor(0, 0, 320 << 2, 240 << 2)
    cl.set_other_modes(rs.OtherModes(cycle_type=0))
    shade = ((rs.MULADD_ZERO, rs.MULSUB_ZERO, rs.MUL_ZERO, rs.RGB_SHADE), (rs.A_ZERO, rs.A_ZERO, rs.AMUL_ZERO, rs.A_SHADE))
    cl.set_combine(shade[0], shade[1])
    xy = np.stack([rng.integers(0, 320 * 4, size=(40, 3)), rng.integers(0, 240 * 4, size=(40, 3))], axis=-1)
    cl.extend_array(rs.triangle_commands(xy, rgba=rng.integers(0, 256, size=(40, 3, 4)).astype(np.float64)))
    cl.sync_pipe()
    cl.set_other_modes(rs.OtherModes(cycle_type=3))
    cl.set_color_image(0x180000, 160, rs.SIZ_16, rs.FMT_RGBA)
    cl.set_scissor(0, 0, 160 << 2, 100 << 2)
    for _ in range(30):
        cl.set_fill_color(int(rng.integers(0, 1 << 32, dtype=np.int64)))
        fx, fy = int(rng.integers(0, 600)), int(rng.integers(0, 380))
        cl.fill_rectangle(fx, fy, fx + int(rng.integers(0, 200)), fy + int(rng.integers(0, 100)))
    cl.sync_full()
    rdram = rng.integers(0, 1 << 32, size=rs.RDRAM_SIZE // 4, dtype=np.uint64).astype(np.uint32)
    return rs.Stream(name=f"fill{seed}", seed=seed, frames=[rs.Frame(cl.words())], rdram_init=rdram,
                     hidden_init=int(rng.integers(0, 4)))


@pytest.mark.parametrize("seed,many", [(1, False), (2, False), (3, True)])
def test_fill_passes_match_oracle(seed, many):
    s = fill_stream(seed, many)
    print(f"[case] {s.name} many={many}", flush=True)
    total, lines = harness.compare(harness.run_gpu(s), harness.run_oracle(s))
    assert total == 0, f"{s.name}: " + "; ".join(lines)


def test_fill_passes_batched_recording():
    import rdp_mi
    streams = [fill_stream(10 + i) for i in range(6)]
    print(f"[case] batched recording of {len(streams)} fill streams", flush=True)
    procs = [rdp_mi.CommandProcessor(rs.RDRAM_SIZE) for _ in streams]
    for p, s in zip(procs, streams):
        p.hold_submissions()
        p.write_rdram(s.rdram_init)
        p.write_hidden_rdram(np.full(rs.RDRAM_SIZE // 2, s.hidden_init, np.uint8))
        p.enqueue_commands(s.frames[0].words)
    rec = rdp_mi.Recording(procs, batched=True)
    assert rec.passes >= len(streams)
    rec.replay(1)
    outs = [(p.read_rdram(), p.read_hidden_rdram(), p.read_tmem()) for p in procs]
    rec.close()
    for p in procs:
        p.close()
    for out, s in zip(outs, streams):
        total, lines = harness.compare(out, harness.run_oracle(s))
        assert total == 0, f"{s.name}: " + "; ".join(lines)
