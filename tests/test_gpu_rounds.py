"""Order between a round's TMEM replay and the previous round's raster.

In a submission of 8 or more processors the TMEM replay of round r runs on a
side stream; it starts before round r-1's raster only when the host shows
that no upload of round r reads RDRAM that the same processor's round r-1
pass can write.  These streams render a texture in pass A (fill rectangles
over random RDRAM) and then, in pass B, load a texture either from that
image (render to texture: the replay must wait) or from elsewhere (it may
start early), and draw textured rectangles with it.  Every processor's
RDRAM, hidden RDRAM and TMEM must match the oracle byte for byte."""
import numpy as np
import pytest

import harness
import rdp_streams as rs

pytestmark = pytest.mark.gpu

SRC, OTHER, DST = 0x200000, 0x280000, 0x300000
TEXEL0 = ((rs.MULADD_ZERO, rs.MULSUB_ZERO, rs.MUL_ZERO, rs.RGB_TEXEL0), (rs.A_ZERO, rs.A_ZERO, rs.AMUL_ZERO, rs.A_TEXEL0))


def texture_stream(seed: int, from_pass_a: bool) -> rs.Stream:
    rng = np.random.default_rng(seed)
    cl = rs.CommandList()
    # Pass A: the texture image, 32x32 RGBA5551, drawn with fill rectangles.
    cl.set_other_modes(rs.OtherModes(cycle_type=3))
    cl.set_color_image(SRC, 32, rs.SIZ_16, rs.FMT_RGBA)
    cl.set_scissor(0, 0, 32 << 2, 32 << 2)
    for _ in range(12):
        cl.set_fill_color(int(rng.integers(0, 1 << 32, dtype=np.int64)))
        x, y = int(rng.integers(0, 28)), int(rng.integers(0, 28))
        cl.fill_rectangle(x << 2, y << 2, (x + int(rng.integers(1, 12))) << 2, (y + int(rng.integers(1, 12))) << 2)
    cl.sync_pipe()
    # Pass B: textured rectangles into another image, texture loaded from
    # pass A's image or from untouched memory.
    cl.set_color_image(DST, 64, rs.SIZ_16, rs.FMT_RGBA)
    cl.set_scissor(0, 0, 64 << 2, 64 << 2)
    cl.set_other_modes(rs.OtherModes(cycle_type=0))
    cl.set_combine(*TEXEL0)
    cl.set_texture_image(SRC if from_pass_a else OTHER, 32, rs.SIZ_16, rs.FMT_RGBA)
    cl.set_tile(7, rs.FMT_RGBA, rs.SIZ_16, 0, 64)
    cl.sync_load()
    cl.load_tile(7, 0, 0, 31 << 2, 31 << 2)
    cl.sync_tile()
    cl.set_tile(0, rs.FMT_RGBA, rs.SIZ_16, 0, 64, mask_s=5, mask_t=5)
    cl.set_tile_size(0, 0, 0, 31 << 2, 31 << 2)
    for _ in range(6):
        x, y = int(rng.integers(0, 48)), int(rng.integers(0, 48))
        cl.texture_rectangle(0, x << 2, y << 2, (x + int(rng.integers(4, 30))) << 2, (y + int(rng.integers(4, 30))) << 2,
                             int(rng.integers(0, 32)) << 5, int(rng.integers(0, 32)) << 5, 1 << 10, 1 << 10)
    cl.sync_full()
    rdram = rng.integers(0, 1 << 32, size=rs.RDRAM_SIZE // 4, dtype=np.uint64).astype(np.uint32)
    return rs.Stream(name=f"rtt{seed}:{'A' if from_pass_a else 'other'}", seed=seed, frames=[rs.Frame(cl.words())],
                     rdram_init=rdram, hidden_init=3)


@pytest.mark.parametrize("sources", ["pass_a", "other", "mixed"])
def test_tmem_replay_order_across_rounds(sources):
    import rdp_mi
    n = 10
    flags = {"pass_a": [True] * n, "other": [False] * n, "mixed": [i % 3 == 0 for i in range(n)]}[sources]
    streams = [texture_stream(100 + i, f) for i, f in enumerate(flags)]
    print(f"[case] batched recording of {n} render-to-texture streams, sources={sources}", flush=True)
    procs = [rdp_mi.CommandProcessor(rs.RDRAM_SIZE) for _ in streams]
    for p, s in zip(procs, streams):
        p.hold_submissions()
        p.write_rdram(s.rdram_init)
        p.write_hidden_rdram(np.full(rs.RDRAM_SIZE // 2, s.hidden_init, np.uint8))
        p.enqueue_commands(s.frames[0].words)
    rec = rdp_mi.Recording(procs, batched=True)
    assert rec.rounds == 2 and rec.passes == 2 * n
    rec.replay(1)
    outs = [(p.read_rdram(), p.read_hidden_rdram(), p.read_tmem()) for p in procs]
    rec.close()
    for p in procs:
        p.close()
    for out, s in zip(outs, streams):
        total, lines = harness.compare(out, harness.run_oracle(s))
        assert total == 0, f"{s.name}: " + "; ".join(lines)


def test_tmem_replay_order_through_server():
    """The served form batches the same frames: 12 instances, requests queued
    together and a batching window that holds the first batch until every
    frame has decoded, so one batch carries every instance's two passes
    (rounds of 12 passes: the side-stream TMEM replay runs)."""
    import rdp_mi
    streams = [texture_stream(200 + i, i % 2 == 0) for i in range(12)]
    print(f"[case] server batch of {len(streams)} render-to-texture frames", flush=True)
    procs = [rdp_mi.CommandProcessor(rs.RDRAM_SIZE) for _ in streams]
    for p, s in zip(procs, streams):
        p.write_rdram(s.rdram_init)
        p.write_hidden_rdram(np.full(rs.RDRAM_SIZE // 2, s.hidden_init, np.uint8))
    with rdp_mi.Server(procs, streams=1) as srv:
        srv.set_batch_window(10e6)
        for inst, s in enumerate(streams):
            srv.submit(inst, s.frames[0].words, arrival_us=0.0)
        srv.start()
        srv.drain()
        assert srv.results()["ticket"].size == len(streams)
        st = srv.stats()
        assert st["batches"] == 1 and st["max_batch_frames"] == len(streams), st
    outs = [(p.read_rdram(), p.read_hidden_rdram(), p.read_tmem()) for p in procs]
    for p in procs:
        p.close()
    for out, s in zip(outs, streams):
        total, lines = harness.compare(out, harness.run_oracle(s))
        assert total == 0, f"{s.name}: " + "; ".join(lines)
