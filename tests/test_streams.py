"""Host-side stream generation: deterministic, well-formed command words."""
import numpy as np
import pytest

import rdp_streams as rs


def walk(words):
    """Split a word stream into commands using the RDP length table."""
    i, ops = 0, []
    while i < len(words):
        op = (int(words[i]) >> 24) & 63
        n = int(rs.COMMAND_LENGTHS[op])
        assert n > 0, f"undefined opcode {op:#x} at word {i}"
        ops.append(op)
        i += n
    assert i == len(words)
    return ops


def test_command_lengths_match_oracle(oracle_mod):
    for op in range(64):
        assert oracle_mod.command_length(op << 24) == int(rs.COMMAND_LENGTHS[op]) or op == 4


def test_batch_streams_are_deterministic_and_well_formed():
    a = rs.batch_streams(6)
    b = rs.batch_streams(6)
    for x, y in zip(a, b):
        assert x.name == y.name and x.num_triangles == y.num_triangles
        assert np.array_equal(x.frames[0].words, y.frames[0].words)
        ops = walk(x.frames[0].words)
        assert ops[-1] == rs.OP_SYNC_FULL
        # num_triangles counts triangles and texture rectangles (clears excluded)
        assert sum(1 for o in ops if 0x08 <= o <= 0x0F or o in (0x24, 0x25)) == x.num_triangles


def test_fuzz_streams_are_deterministic_and_well_formed():
    for seed in range(8):
        a, b = rs.fuzz_stream(seed), rs.fuzz_stream(seed)
        assert np.array_equal(a.frames[0].words, b.frames[0].words)
        assert np.array_equal(a.rdram_init, b.rdram_init)
        walk(a.frames[0].words)


def test_triangle_setup_matches_reference_edge_rules():
    # One right-major triangle with a fractional top vertex: the slope is
    # rounded away from zero and masked with ~7, the start moved back by the
    # sub-scanline offset (triangle_converter.cpp:129-200).
    xy = np.array([[[40, 9], [200, 100], [10, 180]]])
    w = rs.triangle_commands(xy)[0]
    assert (w[0] >> 24) & 63 == rs.OP_FILL_TRIANGLE
    yh, ym, yl = w[0] & 0x3FFF, (w[1] >> 16) & 0x3FFF, w[1] & 0x3FFF
    assert (yl, ym, yh) == (9, 100, 180)
    dxdy_h = np.int32(np.uint32(w[5]))
    exact = ((10 - 40) << 16) / (180 - 9)
    assert abs(int(dxdy_h) - exact) < 8 and int(dxdy_h) & 7 == 0
    assert int(np.int32(np.uint32(w[4]))) == (40 << 14) - (int(dxdy_h) >> 2) * (9 & 3)


def test_zero_area_triangles_are_dropped():
    xy = np.array([[[0, 0], [10, 10], [20, 20]], [[0, 0], [40, 0], [0, 40]]])
    assert rs.triangle_commands(xy).shape[0] == 1


@pytest.mark.parametrize("mode", rs.DENSE_MODES)
def test_dense_streams_cover_every_pixel_once_per_layer(oracle_mod, mode):
    """The per-fragment cost calibration counts frame 1 of a dense stream as
    layers x width x height fragments, all in the stream's mode."""
    s = rs.dense_stream(mode, width=160, height=120, layers=4)
    o = oracle_mod.Oracle(rs.RDRAM_SIZE)
    if s.rdram_init is not None:
        o.rdram[:] = s.rdram_init
    o.hidden[:] = s.hidden_init
    o.process(s.frames[0].words)
    o.flush()
    before = o.fragments_by_mode()
    o.process(s.frames[1].words)
    o.flush()
    after = o.fragments_by_mode()
    o.close()
    delta = {k: after[k] - before[k] for k in after}
    assert delta[mode] == s.num_triangles * s.width * s.height
    assert sum(delta.values()) == delta[mode]
