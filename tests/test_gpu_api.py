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


def test_upscaling_flag_is_rejected():
    import rdp_mi
    with pytest.raises(rdp_mi.RDPError):
        rdp_mi.CommandProcessor(rs.RDRAM_SIZE, flags=1 << 2)


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
