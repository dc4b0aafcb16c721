"""Continuous-batching server: every frame result is the same alone, in the
mix and on the oracle.

Each simulated emulator instance owns one processor and sends its command
stream as a sequence of frame requests (the stream split after every
SyncFull).  Requests of all instances arrive interleaved, so the server's
batches mix frames of different instances, sizes and formats, and a frame of
one instance is admitted while frames of others are still on the GPU.  After
the server drains, each instance's RDRAM, hidden RDRAM and TMEM must equal
(a) the oracle run over the whole stream and (b) the same stream rendered
alone through the plain entry points.  Fuzz streams carry state from frame
to frame (blending reads memory, TMEM persists), so a frame rendered out of
order, twice, or with another instance's state would show in the final
bytes.
"""
import numpy as np
import pytest

from harness import compare, run_gpu, run_oracle, rs

pytestmark = pytest.mark.gpu


def _instances(n_fuzz=10, n_batch=6):
    streams = [rs.fuzz_stream(0x5E2F000 + i, segments=16) for i in range(n_fuzz)]
    # Batch-form frames (shaded, textured, 2-cycle) repeated three times.
    for i in range(n_batch):
        s = rs.batch_stream(i, max_tris=600)
        s.frames = [rs.Frame(s.frames[0].words) for _ in range(3)]
        streams.append(s)
    return streams


def _requests(streams):
    reqs = []
    for inst, s in enumerate(streams):
        for f in s.frames:
            reqs.extend((inst, part) for part in rs.split_at_sync_full(f.words))
    return reqs


def _init(p, s):
    p.write_rdram(s.rdram_init if s.rdram_init is not None else np.zeros(rs.RDRAM_SIZE // 4, np.uint32))
    p.write_hidden_rdram(np.full(rs.RDRAM_SIZE // 2, s.hidden_init, np.uint8))


def _serve(instance_streams, spacing_us, **server_kw):
    import rdp_mi
    procs = [rdp_mi.CommandProcessor(rs.RDRAM_SIZE) for _ in instance_streams]
    for p, s in zip(procs, instance_streams):
        _init(p, s)
    reqs = _requests(instance_streams)
    # Interleave the instances: request k of every instance before request k+1.
    rank, seen = [], {}
    for inst, _ in reqs:
        rank.append(seen.get(inst, 0))
        seen[inst] = rank[-1] + 1
    order = sorted(range(len(reqs)), key=lambda k: (rank[k], reqs[k][0]))
    tickets = {}
    with rdp_mi.Server(procs, **server_kw) as srv:
        for slot, k in enumerate(order):
            inst, words = reqs[k]
            tickets[srv.submit(inst, words, arrival_us=slot * spacing_us)] = inst
        srv.start()
        srv.drain()
        res = srv.results()
    outs = [(p.read_rdram(), p.read_hidden_rdram(), p.read_tmem()) for p in procs]
    for p in procs:
        p.close()
    return outs, res, tickets


@pytest.mark.parametrize("spacing_us,kw", [
    (0.0, {}),                                   # everything queued at once
    (40.0, {}),                                  # arrivals while batches run
    (10.0, {"max_batch": 3}),                    # small slots
    (10.0, {"hbm_budget_bytes": 1}),             # budget admits one frame at a time
    (0.0, {"decode_threads": 1}),
    (20.0, {"streams": 1}),
    (20.0, {"streams": 4, "decode_threads": 3}),
])
def test_serving_matches_alone_and_oracle(spacing_us, kw):
    streams = _instances()
    print(f"serving case spacing={spacing_us}us {kw}: {len(streams)} instances", flush=True)
    outs, res, tickets = _serve(streams, spacing_us, **kw)

    # Every request completed once, in arrival order per instance, with
    # ordered timestamps.
    assert sorted(res["ticket"].tolist()) == sorted(tickets)
    assert all(tickets[int(t)] == int(i) for t, i in zip(res["ticket"], res["instance"]))
    assert np.all(res["admit_us"] >= res["arrival_us"] - 1e-6)
    assert np.all(res["complete_us"] >= res["admit_us"])
    for inst in range(len(streams)):
        mine = res["ticket"][res["instance"] == inst]
        assert np.all(np.diff(mine.astype(np.int64)) > 0), f"instance {inst} completed out of order"

    for inst, s in enumerate(streams):
        alone = run_gpu(s)
        n_alone, lines = compare(outs[inst], alone)
        assert n_alone == 0, f"{s.name} in the mix differs from alone: " + "; ".join(lines)
        ref = run_oracle(s)
        n_ref, lines = compare(outs[inst], ref)
        assert n_ref == 0, f"{s.name} differs from the oracle: " + "; ".join(lines)


def test_server_rejects_bad_instance():
    import rdp_mi
    p = rdp_mi.CommandProcessor(rs.RDRAM_SIZE)
    with rdp_mi.Server([p]) as srv:
        with pytest.raises(rdp_mi.RDPError):
            srv.submit(1, np.zeros(2, np.uint32))
        srv.start()
        srv.drain()
        assert srv.results()["ticket"].size == 0
    p.close()


def test_server_reset_keeps_serving_exact():
    """A server reset between two halves of every instance's stream leaves
    the results identical to the oracle's run of the whole stream."""
    import rdp_mi
    streams = _instances(n_fuzz=6, n_batch=3)
    print(f"serving case reset: {len(streams)} instances", flush=True)
    procs = [rdp_mi.CommandProcessor(rs.RDRAM_SIZE) for _ in streams]
    for p, s in zip(procs, streams):
        _init(p, s)
    parts = [[] for _ in streams]
    for inst, words in _requests(streams):
        parts[inst].append(words)
    with rdp_mi.Server(procs) as srv:
        srv.start()
        for half in (0, 1):
            for inst, frames in enumerate(parts):
                k = len(frames) // 2
                for j, words in enumerate(frames[:k] if half == 0 else frames[k:]):
                    srv.submit(inst, words, arrival_us=j * 20.0)
            srv.drain()
            res = srv.results()
            assert res["ticket"].size == sum(len(f) // 2 if half == 0 else len(f) - len(f) // 2 for f in parts)
            srv.reset(0.0)
            assert srv.results()["ticket"].size == 0
    for p, s in zip(procs, streams):
        n, lines = compare((p.read_rdram(), p.read_hidden_rdram(), p.read_tmem()), run_oracle(s))
        assert n == 0, f"{s.name}: " + "; ".join(lines)
        p.close()
