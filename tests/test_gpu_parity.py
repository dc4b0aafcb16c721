"""Parity of the MI355X path with the oracle: RDRAM, hidden RDRAM and TMEM
must match byte for byte after every stream."""
import numpy as np
import pytest

import harness
import rdp_streams as rs

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)


def assert_same(gpu, ref, what):
    total, lines = harness.compare(gpu, ref)
    assert total == 0, f"{what}: " + "; ".join(lines)


@pytest.fixture(scope="module")
def rdp_mi():
    import rdp_mi as mod
    mod.load_library()
    return mod


def test_config0_plain_shaded(rdp_mi):
    s = rs.config0_stream()
    say(f"[case] {s.name} {s.num_triangles} triangles")
    assert_same(harness.run_gpu(s), harness.run_oracle(s), s.name)


def test_config1_textured_two_frames(rdp_mi):
    s = rs.config1_stream(frames=2)
    say(f"[case] {s.name} 2 frames, {s.num_triangles} triangles")
    assert_same(harness.run_gpu(s), harness.run_oracle(s), s.name)


def test_config2_two_cycle_with_vi(rdp_mi):
    s = rs.config2_stream(frames=3)
    say(f"[case] {s.name} 3 frames, {s.num_triangles} primitives, VI scanout per frame")
    gimg, oimg = [], []
    assert_same(harness.run_gpu(s, scanouts=gimg), harness.run_oracle(s, scanouts=oimg), s.name)
    assert len(gimg) == len(oimg) == 3
    for k, (g, o) in enumerate(zip(gimg, oimg)):
        assert g.shape == o.shape == (240, 640), (k, g.shape, o.shape)
        assert np.array_equal(g, o), f"VI frame {k}: {int(np.count_nonzero(g != o))} pixels differ"


@pytest.mark.parametrize("start", [0, 8, 16, 24])
def test_batch_streams_immediate(rdp_mi, start):
    for i in range(start, start + 8):
        s = rs.batch_stream(i)
        say(f"[case] {s.name} {s.num_triangles} triangles")
        assert_same(harness.run_gpu(s), harness.run_oracle(s), s.name)


@pytest.mark.parametrize("seed", range(0, 96, 8))
def test_fuzz_streams(rdp_mi, seed):
    for k in range(seed, seed + 8):
        s = rs.fuzz_stream(k)
        say(f"[case] {s.name}")
        assert_same(harness.run_gpu(s), harness.run_oracle(s), s.name)


def _procs_for(rdp_mi, streams):
    procs = []
    for s in streams:
        p = rdp_mi.CommandProcessor(rs.RDRAM_SIZE)
        p.hold_submissions()  # every pass stays pending for the recording or batched flush
        p.write_rdram(s.rdram_init if s.rdram_init is not None else np.zeros(rs.RDRAM_SIZE // 4, np.uint32))
        p.write_hidden_rdram(np.full(rs.RDRAM_SIZE // 2, s.hidden_init, np.uint8))
        for f in s.frames:
            p.enqueue_commands(f.words)
        procs.append(p)
    return procs


def _state(p):
    return p.read_rdram(), p.read_hidden_rdram(), p.read_tmem()


@pytest.mark.parametrize("batched", [True, False])
def test_recorded_fuzz_batch(rdp_mi, batched):
    streams = [rs.fuzz_stream(1000 + i) for i in range(12)]
    say(f"[case] recorded fuzz batch batched={batched}")
    procs = _procs_for(rdp_mi, streams)
    rec = rdp_mi.Recording(procs, batched=batched)
    rec.replay(1, use_graph=False)
    for p, s in zip(procs, streams):
        assert_same(_state(p), harness.run_oracle(s), s.name)
    rec.close()
    for p in procs:
        p.close()


def test_graph_replays_match_eager(rdp_mi):
    """Frames that start with a clear are idempotent: every graph replay,
    with eager submissions and synchronisation between replays, must give
    the oracle's result."""
    streams = [rs.batch_stream(i) for i in (1, 4, 9, 17)]
    refs = [harness.run_oracle(s) for s in streams]
    say("[case] graph replays")
    procs = _procs_for(rdp_mi, streams)
    rec = rdp_mi.Recording(procs, batched=True)
    side = rdp_mi.CommandProcessor(rs.RDRAM_SIZE)
    side_stream = rs.config0_stream(num_triangles=32)
    side_ref = harness.run_oracle(side_stream)
    for it in range(3):
        rec.replay(1, use_graph=True)
        for p, s, r in zip(procs, streams, refs):
            assert_same(_state(p), r, f"{s.name} replay {it}")
        assert_same(harness.run_gpu(side_stream, proc=side), side_ref, f"eager between replays {it}")
    rec.replay(2, use_graph=False)
    for p, s, r in zip(procs, streams, refs):
        assert_same(_state(p), r, f"{s.name} direct replay")
    rec.close()
    side.close()
    for p in procs:
        p.close()


def test_batch_flush_many_processors(rdp_mi):
    streams = [rs.batch_stream(i, max_tris=600) for i in range(20)]
    say("[case] batch_flush of 20 processors")
    procs = []
    for s in streams:
        p = rdp_mi.CommandProcessor(rs.RDRAM_SIZE)
        p.hold_submissions()
        p.write_rdram(s.rdram_init if s.rdram_init is not None else np.zeros(rs.RDRAM_SIZE // 4, np.uint32))
        p.write_hidden_rdram(np.full(rs.RDRAM_SIZE // 2, s.hidden_init, np.uint8))
        procs.append(p)
    # Interleave enqueues so every processor has pending passes, then submit together.
    for p, s in zip(procs, streams):
        p.enqueue_commands(s.frames[0].words)
    rdp_mi.batch_flush(procs)
    for p, s in zip(procs, streams):
        p.idle()
        assert_same(_state(p), harness.run_oracle(s), s.name)
        p.close()


@pytest.mark.parametrize("xcd_chunk,min_class,tile_size,persistent",
                         [(-1, 0, 8, 0), (0, 1, 8, 0), (32, 2, 8, 0), (128, 0, 8, 0), (64, 1, 8, 0),
                          (-1, 0, 16, 0), (0, 1, 16, 0), (32, 2, 16, 0), (128, 0, 16, 0), (64, 1, 16, 0), (64, 2, 16, 0),
                          (64, 0, 8, 1), (64, 1, 8, 1), (64, 2, 8, 1)])
def test_launch_shapes_bit_exact(rdp_mi, xcd_chunk, min_class, tile_size, persistent):
    """Every raster launch shape the autotuner can pick gives the oracle's bytes:
    batch streams of all three workloads plus random-state streams, submitted
    together."""
    streams = [rs.batch_stream(i, max_tris=800) for i in range(6)] + [rs.fuzz_stream(2000 + i) for i in range(6)]
    say(f"[case] launch shape xcd_chunk={xcd_chunk} min_class={min_class} tile_size={tile_size} persistent={persistent}")
    rdp_mi.set_launch_config(xcd_chunk, min_class, tile_size, bool(persistent))
    try:
        procs = _procs_for(rdp_mi, streams)
        rdp_mi.batch_flush(procs)
        for p, s in zip(procs, streams):
            p.idle()
            assert_same(_state(p), harness.run_oracle(s), f"{s.name} ({xcd_chunk}, {min_class}, {tile_size}, {persistent})")
            p.close()
    finally:
        rdp_mi.set_launch_config(64, 0, 8)


def test_autotuned_recording_bit_exact(rdp_mi):
    streams = [rs.batch_stream(i) for i in (2, 5, 11, 23)]
    refs = [harness.run_oracle(s) for s in streams]
    say("[case] autotuned recording")
    procs = _procs_for(rdp_mi, streams)
    rec = rdp_mi.Recording(procs, batched=True)
    tuned = rec.autotune(1)
    say(f"[case] autotune picked {tuned}")
    assert tuned["min_class"] in (0, 1, 2) and tuned["xcd_chunk"] in (-1, 0, 32, 64, 128)
    assert tuned["tile_size"] in (8, 16)
    rec.replay(2, use_graph=True)
    for p, s, r in zip(procs, streams, refs):
        assert_same(_state(p), r, f"{s.name} after autotune")
    rec.close()
    for p in procs:
        p.close()


def test_trial_runs_leave_state_unchanged(rdp_mi):
    """Autotune and profile execute the recording many times; random-state
    streams are not idempotent, so only a saved and restored state lets one
    replay afterwards give the oracle's result."""
    streams = [rs.fuzz_stream(3000 + i) for i in range(6)]
    refs = [harness.run_oracle(s) for s in streams]
    say("[case] autotune and profile on non-idempotent streams")
    procs = _procs_for(rdp_mi, [rs.Stream(s.name, s.seed, [], s.rdram_init, s.hidden_init) for s in streams])
    before = [_state(p) for p in procs]
    for p, s in zip(procs, streams):
        for f in s.frames:
            p.enqueue_commands(f.words)
    rec = rdp_mi.Recording(procs, batched=True)
    rec.autotune(1)
    rec.profile(2)
    rec.profile_rounds(1)
    for p, b, s in zip(procs, before, streams):
        assert_same(_state(p), b, f"{s.name} state after trial runs")
    rec.replay(1, use_graph=True)
    for p, s, r in zip(procs, streams, refs):
        assert_same(_state(p), r, f"{s.name} after one replay")
    rec.close()
    for p in procs:
        p.close()


def test_persistent_round_launch_graph_replays(rdp_mi):
    """A recording made under the persistent round launch, replayed three
    times as one captured graph: every replay leaves the oracle's bytes (the
    batch frames clear first, so replays are idempotent)."""
    streams = [rs.batch_stream(i, max_tris=1500) for i in range(10)]
    refs = [harness.run_oracle(s) for s in streams]
    say("[case] persistent round launch, graph replays")
    rdp_mi.set_launch_config(64, 0, 8, True)
    try:
        procs = _procs_for(rdp_mi, streams)
        rec = rdp_mi.Recording(procs, batched=True)
    finally:
        rdp_mi.set_launch_config(64, 0, 8)
    for it in range(3):
        rec.replay(1, use_graph=True)
        for p, s, r in zip(procs, streams, refs):
            assert_same(_state(p), r, f"{s.name} persistent graph replay {it}")
    rec.close()
    for p in procs:
        p.close()
