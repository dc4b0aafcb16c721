"""rdp_mi_batch_flush_noalloc (packed inputs and scratch in the library's
static device array) against rdp_mi_batch_flush: byte-identical RDRAM, hidden
RDRAM and TMEM on two calls in a row, the second with a smaller input, and a
submission beyond the array's limit refused before any launch with its
passes left pending."""
import numpy as np
import pytest

import harness
import rdp_streams as rs

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module")
def rdp_mi():
    import rdp_mi as mod
    mod.load_library()
    return mod


def _procs(rdp_mi, streams):
    procs = []
    for s in streams:
        p = rdp_mi.CommandProcessor(rs.RDRAM_SIZE)
        p.hold_submissions()  # every pass stays pending for the batched flush
        p.write_rdram(s.rdram_init if s.rdram_init is not None else np.zeros(rs.RDRAM_SIZE // 4, np.uint32))
        p.write_hidden_rdram(np.full(rs.RDRAM_SIZE // 2, s.hidden_init, np.uint8))
        for f in s.frames:
            p.enqueue_commands(f.words)
        procs.append(p)
    return procs


def _state(p):
    return p.read_rdram(), p.read_hidden_rdram(), p.read_tmem()


def _same(a, b, what):
    total, lines = harness.compare(a, b)
    assert total == 0, f"{what}: " + "; ".join(lines)


def _check_call(rdp_mi, streams, label, oracle=False):
    ref_procs = _procs(rdp_mi, streams)
    rdp_mi.batch_flush(ref_procs)
    procs = _procs(rdp_mi, streams)
    say(f"[case] {label}: batch_flush_noalloc of {len(streams)} processors")
    assert rdp_mi.batch_flush_noalloc(procs) == 0
    for p, q, s in zip(procs, ref_procs, streams):
        q.idle()
        p.idle()
        got = _state(p)
        _same(got, _state(q), f"{label} {s.name} vs batch_flush")
        if oracle:
            _same(got, harness.run_oracle(s), f"{label} {s.name} vs oracle")
    for p in procs + ref_procs:
        p.close()


def test_noalloc_two_calls_second_smaller(rdp_mi):
    big = [rs.batch_stream(i, max_tris=1500) for i in range(8)] + [rs.fuzz_stream(5100 + i) for i in range(4)]
    small = [rs.batch_stream(40 + i, max_tris=300) for i in range(2)] + [rs.fuzz_stream(5200)]
    _check_call(rdp_mi, big, "call 1 (12 streams)", oracle=True)
    _check_call(rdp_mi, small, "call 2 (3 streams)", oracle=True)


def test_noalloc_beyond_limit_keeps_passes(rdp_mi):
    # Two 640x480 streams of 8192 full-screen triangles: one round whose
    # span records alone exceed the 112 MiB of scratch.
    streams = [rs.config0_stream(seed=0x5244F000 + i, num_triangles=8192, width=640, height=480) for i in range(2)]
    procs = _procs(rdp_mi, streams)
    say("[case] batch_flush_noalloc beyond the static array")
    assert rdp_mi.batch_flush_noalloc(procs) == rdp_mi.ERROR_LIMIT
    # The passes are still pending: the allocating entry point renders them.
    rdp_mi.batch_flush(procs)
    ref_procs = _procs(rdp_mi, streams)
    rdp_mi.batch_flush(ref_procs)
    for p, q, s in zip(procs, ref_procs, streams):
        p.idle()
        q.idle()
        _same(_state(p), _state(q), f"{s.name} after a refused call")
    for p in procs + ref_procs:
        p.close()
    # A later call that fits still works after the refusal.
    _check_call(rdp_mi, [rs.batch_stream(3, max_tris=400)], "call after the refusal")
