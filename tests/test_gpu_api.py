"""CommandProcessor semantics on the device: timelines, idle, memory access,
sizes, validation reports and error handling (rdp_device.hpp:111-198)."""
import numpy as np
import pytest

import harness
import rdp_streams as rs

pytestmark = pytest.mark.gpu


@pytest.fixture
def proc():
    import rdp_mi
    p = rdp_mi.CommandProcessor(rs.RDRAM_SIZE)
    yield p
    p.close()


def test_fresh_processor_state(proc):
    assert proc.device_is_supported()
    assert not proc.read_rdram().any()
    assert (proc.read_hidden_rdram() == 3).all()
    assert not proc.read_tmem().any()


def test_rdram_round_trip(proc):
    data = np.random.default_rng(1).integers(0, 1 << 32, size=rs.RDRAM_SIZE // 4, dtype=np.uint64).astype(np.uint32)
    proc.write_rdram(data)
    assert np.array_equal(proc.read_rdram(), data)
    hid = np.random.default_rng(2).integers(0, 4, size=rs.RDRAM_SIZE // 2).astype(np.uint8)
    proc.write_hidden_rdram(hid)
    assert np.array_equal(proc.read_hidden_rdram(), hid)


def test_timelines_are_monotonic_and_complete(proc):
    s = rs.config0_stream(num_triangles=64)
    words = s.frames[0].words
    proc.write_hidden_rdram(np.full(rs.RDRAM_SIZE // 2, s.hidden_init, np.uint8))
    half = len(words) // 2
    # Split on a command boundary.
    i = 0
    while i < half:
        i += int(rs.COMMAND_LENGTHS[(int(words[i]) >> 24) & 63])
    proc.enqueue_commands(words[:i])
    t1 = proc.signal_timeline()
    proc.enqueue_commands(words[i:])
    t2 = proc.signal_timeline()
    assert t2 > t1
    proc.wait_for_timeline(t2)
    proc.wait_for_timeline(t1)
    ref = harness.run_oracle(s)
    total, lines = harness.compare((proc.read_rdram(), proc.read_hidden_rdram(), proc.read_tmem()), ref)
    assert total == 0, lines


def test_word_by_word_enqueue_matches_bulk(proc):
    s = rs.batch_stream(3)
    words = s.frames[0].words
    proc.write_rdram(s.rdram_init if s.rdram_init is not None else np.zeros(rs.RDRAM_SIZE // 4, np.uint32))
    proc.write_hidden_rdram(np.full(rs.RDRAM_SIZE // 2, s.hidden_init, np.uint8))
    i = 0
    while i < len(words):
        n = int(rs.COMMAND_LENGTHS[(int(words[i]) >> 24) & 63])
        proc.enqueue_command(words[i:i + n])
        i += n
    proc.idle()
    ref = harness.run_oracle(s)
    total, lines = harness.compare((proc.read_rdram(), proc.read_hidden_rdram(), proc.read_tmem()), ref)
    assert total == 0, lines


def test_primitive_count_counts_draw_commands(proc):
    s = rs.config0_stream(num_triangles=50)
    proc.enqueue_commands(s.frames[0].words)
    proc.idle()
    assert proc.primitive_count == s.num_triangles + 1  # plus the depth clear rectangle


def test_validation_reports_fill_mode_depth_test(proc):
    proc.set_validation_callback()
    cl = rs.CommandList()
    cl.set_color_image(rs.COLOR_ADDR, 64, rs.SIZ_16)
    cl.set_scissor(0, 0, 64 << 2, 64 << 2)
    cl.set_other_modes(rs.OtherModes(cycle_type=3, depth_test=True))
    cl.fill_rectangle(0, 0, 16 << 2, 16 << 2)
    cl.sync_full()
    proc.enqueue_commands(cl.words())
    proc.idle()
    assert any("depth test" in m for _, m in proc.validation_errors)


def test_upscaling_with_host_coherent_rdram_is_rejected():
    """Upscaling needs device-resident RDRAM (RDP_MI_ERROR_UNSUPPORTED)."""
    import rdp_mi
    host = np.zeros(rs.RDRAM_SIZE // 4, dtype=np.uint32)
    with pytest.raises(rdp_mi.RDPError):
        rdp_mi.CommandProcessor(rs.RDRAM_SIZE, flags=rdp_mi.FLAG_UPSCALING[2], host_rdram=host)


def test_small_rdram_wraps_like_the_oracle():
    import rdp_mi
    size = 4 << 20
    s = rs.config0_stream(num_triangles=64)
    p = rdp_mi.CommandProcessor(size)
    p.write_hidden_rdram(np.full(size // 2, s.hidden_init, np.uint8))
    p.enqueue_commands(s.frames[0].words)
    p.idle()
    import oracle
    o = oracle.Oracle(size)
    o.hidden[:] = s.hidden_init
    o.process(s.frames[0].words)
    o.flush()
    assert np.array_equal(p.read_rdram(), o.rdram)
    assert np.array_equal(p.read_hidden_rdram(), o.hidden)
    o.close()
    p.close()


def _prepare(proc, s):
    proc.write_rdram(s.rdram_init if s.rdram_init is not None else np.zeros(rs.RDRAM_SIZE // 4, np.uint32))
    proc.write_hidden_rdram(np.full(rs.RDRAM_SIZE // 2, s.hidden_init, np.uint8))


def test_wait_blocks_until_a_later_signal(proc):
    """wait_for_timeline on a value not yet signalled blocks until another
    thread's signal_timeline reaches it (worker_thread.hpp:69-73)."""
    import threading
    s = rs.config0_stream(num_triangles=64)
    _prepare(proc, s)
    target = proc.signal_timeline() + 1
    proc.wait_for_timeline(target - 1)
    done = threading.Event()
    err = []

    def waiter():
        try:
            proc.wait_for_timeline(target, timeout_us=30_000_000)
        except Exception as e:  # noqa: BLE001
            err.append(e)
        done.set()

    t = threading.Thread(target=waiter)
    t.start()
    assert not done.wait(0.3), "wait_for_timeline returned before its value was signalled"
    proc.enqueue_commands(s.frames[0].words)
    assert proc.signal_timeline() == target
    assert done.wait(30), "wait_for_timeline did not return after the signal"
    t.join()
    assert not err, err
    assert proc.completed_timeline >= target
    total, lines = harness.compare((proc.read_rdram(), proc.read_hidden_rdram(), proc.read_tmem()),
                                   harness.run_oracle(s))
    assert total == 0, lines


def test_wait_with_timeout_reports_an_unsignalled_value(proc):
    import rdp_mi  # noqa: F401
    v = proc.signal_timeline()
    proc.wait_for_timeline(v, timeout_us=10_000_000)
    with pytest.raises(TimeoutError):
        proc.wait_for_timeline(v + 1, timeout_us=20_000)


def test_direct_commands_follow_the_ring_in_order(proc):
    """enqueue_command_direct decodes after everything already in the ring."""
    s = rs.batch_stream(5)
    _prepare(proc, s)
    words = s.frames[0].words
    i, half = 0, len(words) // 2
    while i < half:
        i += int(rs.COMMAND_LENGTHS[(int(words[i]) >> 24) & 63])
    proc.enqueue_commands(words[:i])
    j = i
    import ctypes
    lib = proc._lib
    while j < len(words):
        n = int(rs.COMMAND_LENGTHS[(int(words[j]) >> 24) & 63])
        w = np.ascontiguousarray(words[j:j + n], dtype=np.uint32)
        lib.rdp_mi_enqueue_command_direct(proc.handle, ctypes.c_uint(n), w.ctypes.data)
        j += n
    proc.idle()
    total, lines = harness.compare((proc.read_rdram(), proc.read_hidden_rdram(), proc.read_tmem()),
                                   harness.run_oracle(s))
    assert total == 0, lines


def test_processors_on_separate_threads():
    """Each processor has its own lock and ring: four threads drive four
    processors at once and every result equals the oracle's."""
    import threading
    import rdp_mi
    streams = [rs.batch_stream(k) for k in (1, 4, 9, 12)]
    procs = [rdp_mi.CommandProcessor(rs.RDRAM_SIZE) for _ in streams]
    for p, s in zip(procs, streams):
        _prepare(p, s)
    errs = []

    def drive(p, s):
        try:
            for f in s.frames:
                w = f.words
                k = 0
                while k < len(w):  # in several pieces, with signals between
                    e = min(len(w), k + 4096)
                    while e < len(w) and rs.COMMAND_LENGTHS[(int(w[e]) >> 24) & 63] == 0:
                        e += 1
                    k += p.enqueue_commands(w[k:e])
                    p.signal_timeline()
            p.idle()
        except Exception as ex:  # noqa: BLE001
            errs.append(ex)

    threads = [threading.Thread(target=drive, args=(p, s)) for p, s in zip(procs, streams)]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    assert not errs, errs
    for p, s in zip(procs, streams):
        total, lines = harness.compare((p.read_rdram(), p.read_hidden_rdram(), p.read_tmem()), harness.run_oracle(s))
        assert total == 0, lines
        p.close()


def test_async_scanout_buffer_matches_scanout_sync():
    """scanout_async_buffer returns before the image is ready; after the
    fence it holds the same image as scanout_sync of the same frame."""
    import rdp_mi
    s = rs.config2_stream(frames=2)
    a = rdp_mi.CommandProcessor(rs.RDRAM_SIZE)
    b = rdp_mi.CommandProcessor(rs.RDRAM_SIZE)
    buf = rdp_mi.ScanoutBuffer()
    for p in (a, b):
        _prepare(p, s)
    for f in s.frames:
        for p in (a, b):
            p.enqueue_commands(f.words)
            for r, v in f.vi.items():
                p.set_vi_register(r, v)
        want = a.scanout_sync()
        b.scanout_async(buf)
        got = buf.wait()
        assert want.shape == got.shape and want.size > 0
        assert np.array_equal(want, got)
    buf.close()
    a.close()
    b.close()


def test_held_passes_wait_for_the_recording():
    """With submissions held, passes enqueued long before the recording (well
    past the ring's 500 us idle submission) are all still pending for it, and
    its replay gives the oracle's bytes."""
    import time

    import rdp_mi
    streams = [rs.batch_stream(i, max_tris=400) for i in (1, 4)]
    procs = []
    for s in streams:
        p = rdp_mi.CommandProcessor(rs.RDRAM_SIZE)
        p.hold_submissions()
        p.write_rdram(s.rdram_init if s.rdram_init is not None else np.zeros(rs.RDRAM_SIZE // 4, np.uint32))
        p.write_hidden_rdram(np.full(rs.RDRAM_SIZE // 2, s.hidden_init, np.uint8))
        for f in s.frames:
            p.enqueue_commands(f.words)
        procs.append(p)
    time.sleep(0.02)
    # Nothing reached the device: RDRAM is still the initial image.
    for p, s in zip(procs, streams):
        init = s.rdram_init if s.rdram_init is not None else np.zeros(rs.RDRAM_SIZE // 4, np.uint32)
        assert np.array_equal(p.read_rdram(), init), s.name
    rec = rdp_mi.Recording(procs, batched=True)
    assert rec.passes >= 2 * len(streams)
    rec.replay(1, use_graph=True)
    for p, s in zip(procs, streams):
        total, lines = harness.compare((p.read_rdram(), p.read_hidden_rdram(), p.read_tmem()), harness.run_oracle(s))
        assert total == 0, f"{s.name}: " + "; ".join(lines)
    rec.close()
    for p in procs:
        p.close()
