"""Paths and sizes the other GPU tests leave out: colour/depth aliasing under
shading and depth test/update, the textured workload at its full 60 frames
and the two-cycle workload at its full 30 frames with a VI scanout per frame,
the whole 64-stream batch as one captured graph, and 256 served processors
with the state after every frame checked against the oracle."""
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


def _new_proc(rdp_mi, s):
    p = rdp_mi.CommandProcessor(rs.RDRAM_SIZE)
    p.write_rdram(s.rdram_init if s.rdram_init is not None else np.zeros(rs.RDRAM_SIZE // 4, np.uint32))
    p.write_hidden_rdram(np.full(rs.RDRAM_SIZE // 2, s.hidden_init, np.uint8))
    return p


def _state(p):
    return p.read_rdram(), p.read_hidden_rdram(), p.read_tmem()


@pytest.fixture(scope="module")
def rdp_mi():
    import rdp_mi as mod
    mod.load_library()
    return mod


@pytest.mark.parametrize("partial", [False, True])
@pytest.mark.parametrize("cycle", [0, 1])
@pytest.mark.parametrize("fmt", [rs.FMT_RGBA, rs.FMT_IA])
def test_color_depth_alias(rdp_mi, fmt, cycle, partial):
    for seed in range(6):
        s = rs.alias_stream(0x5100 + 16 * seed + 4 * fmt + cycle, fmt, cycle, partial)
        say(f"[case] {s.name} {s.num_triangles} primitives")
        assert_same(harness.run_gpu(s), harness.run_oracle(s), s.name)


def test_config1_full_sixty_frames(rdp_mi):
    s = rs.config1_stream()
    say(f"[case] {s.name} {len(s.frames)} frames, {s.num_triangles} triangles")
    assert len(s.frames) == 60
    assert_same(harness.run_gpu(s), harness.run_oracle(s), s.name)


def test_config2_full_thirty_frames_with_vi(rdp_mi):
    s = rs.config2_stream()
    say(f"[case] {s.name} {len(s.frames)} frames, {s.num_triangles} primitives, VI scanout per frame")
    assert len(s.frames) == 30
    gimg, oimg = [], []
    assert_same(harness.run_gpu(s, scanouts=gimg), harness.run_oracle(s, scanouts=oimg), s.name)
    assert len(gimg) == len(oimg) == 30
    for k, (g, o) in enumerate(zip(gimg, oimg)):
        assert g.shape == o.shape and g.size > 0, (k, g.shape, o.shape)
        assert np.array_equal(g, o), f"VI frame {k}: {int(np.count_nonzero(g != o))} pixels differ"


def test_batch64_graph_replay(rdp_mi):
    """All 64 streams of the benchmark's batch in one batched recording,
    replayed twice as a captured graph: every stream's state equals the
    oracle's (the frames clear first, so a replay is idempotent)."""
    streams = rs.batch_streams(64)
    say(f"[case] batch of {len(streams)} streams, {sum(s.num_triangles for s in streams)} triangles")
    refs = harness.run_oracle_many(streams)
    procs = [_new_proc(rdp_mi, s) for s in streams]
    for p, s in zip(procs, streams):
        p.hold_submissions()
        for f in s.frames:
            p.enqueue_commands(f.words)
    rec = rdp_mi.Recording(procs, batched=True)
    say(f"[case] recording: {rec.rounds} rounds, {rec.passes} passes")
    assert rec.passes >= 2 * len(streams)
    for it in range(2):
        rec.replay(1, use_graph=True)
        for p, s, r in zip(procs, streams, refs):
            assert_same(_state(p), r, f"{s.name} graph replay {it}")
    rec.close()
    for p in procs:
        p.close()


def test_state_digest_matches_oracle(rdp_mi):
    for s in [rs.fuzz_stream(3100 + i) for i in range(6)] + [rs.batch_stream(7)]:
        p = _new_proc(rdp_mi, s)
        for f in s.frames:
            p.enqueue_commands(f.words)
        p.idle()
        assert p.state_digest() == harness.oracle_frame_digests(s)[-1], s.name
        p.close()


def _multi_frame(seed):
    s = rs.fuzz_stream(seed, segments=10)
    frames = [rs.Frame(w) for w in rs.split_at_sync_full(s.frames[0].words)]
    return rs.Stream(name=f"{s.name}x{len(frames)}", seed=seed, frames=frames, rdram_init=s.rdram_init,
                     hidden_init=s.hidden_init)


def test_serve_256_processors_every_frame_checked(rdp_mi):
    """256 processors, instance i replaying multi-frame stream i mod 64 (state
    carries across frames, so frames are not idempotent).  Requests arrive
    over time and are batched continuously; the device digest of every
    processor's state after every frame must equal the oracle's after the
    same frame."""
    n, distinct = 256, 64
    streams = [_multi_frame(5000 + i) for i in range(distinct)]
    want = [harness.oracle_frame_digests(s) for s in streams]
    say(f"[case] serving {n} processors, {sum(len(s.frames) for s in streams) * n // distinct} frames")
    procs = [_new_proc(rdp_mi, streams[i % distinct]) for i in range(n)]
    rng = np.random.default_rng(8)
    expect = {}
    with rdp_mi.Server(procs, hbm_budget_bytes=2 << 30, max_batch=64) as srv:
        srv.set_digests(True)
        for i in range(n):
            t = 0.0
            for k, f in enumerate(streams[i % distinct].frames):
                t += float(rng.exponential(300.0))
                expect[srv.submit(i, f.words, arrival_us=t)] = want[i % distinct][k]
        srv.start()
        srv.drain()
        d = srv.digests()
        stats = srv.stats()
    say(f"[case] {stats}")
    assert len(d["ticket"]) == len(expect)
    assert not d["status"].any()
    bad = [int(t) for t, g in zip(d["ticket"], d["digest"]) if int(g) != expect[int(t)]]
    assert not bad, f"{len(bad)} of {len(expect)} frames differ, first ticket {bad[0]}"
    assert stats["max_batch_frames"] >= 8
    for p in procs:
        p.close()
