"""Host-coherent RDRAM: a processor created with the caller's RDRAM pointer
(the reference's host-memory RDRAM path, rdp_renderer.cpp:2523-2907,
masked_rdram_resolve.comp:47-70).  The caller's array is RDRAM: it must hold
the oracle's bytes after a wait without any read-back call, caller writes
between frames must reach later passes and scanouts without
end_write_rdram, and caller writes to bytes the raster does not write must
survive the write-back of pages it does write."""
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


def _coherent(rdp_mi, stream):
    host = np.zeros(rs.RDRAM_SIZE // 4, np.uint32)
    if stream.rdram_init is not None:
        host[:] = stream.rdram_init
    p = rdp_mi.CommandProcessor(rs.RDRAM_SIZE, host_rdram=host)
    p.write_hidden_rdram(np.full(rs.RDRAM_SIZE // 2, stream.hidden_init, np.uint8))
    return p, host


def _oracle(stream):
    import oracle
    o = oracle.Oracle(rs.RDRAM_SIZE)
    if stream.rdram_init is not None:
        o.rdram[:] = stream.rdram_init
    o.hidden[:] = stream.hidden_init
    return o


def _assert_state(p, host, o, what):
    o.flush()
    total, lines = harness.compare((host, p.read_hidden_rdram(), p.read_tmem()),
                                   (o.rdram.copy(), o.hidden.copy(), o.tmem.copy()))
    assert total == 0, f"{what}: " + "; ".join(lines)


@pytest.mark.parametrize("kind", ["c0", "textured", "two_cycle", "fuzz"])
def test_caller_memory_holds_the_result(rdp_mi, kind):
    stream = {"c0": rs.config0_stream(num_triangles=128), "textured": rs.config1_stream(frames=2),
              "two_cycle": rs.config2_stream(frames=2, prims_per_frame=400),
              "fuzz": rs.fuzz_stream(0x70C0)}[kind]
    say(f"[case] host-coherent {stream.name}")
    p, host = _coherent(rdp_mi, stream)
    o = _oracle(stream)
    for f in stream.frames:
        p.enqueue_commands(f.words)
        o.process(f.words)
    p.idle()
    _assert_state(p, host, o, stream.name)
    p.close()
    o.close()


def test_caller_writes_between_frames_reach_the_textures(rdp_mi):
    """The caller rewrites the texture images between frames (no
    end_write_rdram): the next frame's loads must see the new texels."""
    stream = rs.config1_stream(frames=3)
    tex = np.nonzero(stream.rdram_init)[0]
    assert tex.size > 0
    say(f"[case] caller texture writes, {tex.size} texture words")
    p, host = _coherent(rdp_mi, stream)
    o = _oracle(stream)
    rng = np.random.default_rng(5)
    for i, f in enumerate(stream.frames):
        if i:
            p.idle()
            o.flush()
            new = rng.integers(0, 1 << 32, size=tex.size, dtype=np.uint64).astype(np.uint32)
            host[tex] = new
            o.rdram[tex] = new
        p.enqueue_commands(f.words)
        o.process(f.words)
    p.idle()
    _assert_state(p, host, o, "after texture rewrites")
    p.close()
    o.close()


def test_caller_writes_in_written_pages_survive(rdp_mi):
    """Bytes after the end of the colour and depth images share a 4 KiB page
    with the last pixels the raster writes; the caller writes them while the
    frame renders, and the masked write-back must keep them."""
    stream = rs.config0_stream(num_triangles=96)
    size = stream.width * stream.height * 2
    spots = []
    for base in (rs.COLOR_ADDR, rs.DEPTH_ADDR):
        end = base + size
        page_end = (end + 4095) & ~4095
        assert page_end - end >= 64
        spots.append(slice(end // 4, page_end // 4))
    say("[case] caller writes beside the images during the frame")
    p, host = _coherent(rdp_mi, stream)
    o = _oracle(stream)
    p.enqueue_commands(stream.frames[0].words)
    p.flush()
    o.process(stream.frames[0].words)
    for k, sl in enumerate(spots):
        pattern = np.full(sl.stop - sl.start, 0xA5C30000 + k, np.uint32)
        host[sl] = pattern
        o.rdram[sl] = pattern
    p.idle()
    _assert_state(p, host, o, "after the masked write-back")
    p.close()
    o.close()


def test_scanout_reads_caller_written_frame(rdp_mi, oracle_mod):
    """A frame the caller writes straight into RDRAM (no RDP drawing) is what
    the VI scans out."""
    stream = rs.config2_stream(frames=1, prims_per_frame=64, width=320, height=240)
    regs = stream.frames[0].vi
    say("[case] scanout of a caller-written frame")
    p, host = _coherent(rdp_mi, stream)
    o = _oracle(stream)
    rng = np.random.default_rng(9)
    words = 320 * 240 * 2 // 4
    img = rng.integers(0, 1 << 32, size=words, dtype=np.uint64).astype(np.uint32)
    host[rs.COLOR_ADDR // 4:rs.COLOR_ADDR // 4 + words] = img
    o.rdram[rs.COLOR_ADDR // 4:rs.COLOR_ADDR // 4 + words] = img
    for r, v in regs.items():
        p.set_vi_register(r, v)
        o.set_vi_register(r, v)
    got = p.scanout_sync()
    ref = o.scanout()
    assert got.shape == ref.shape and got.size > 0
    assert np.array_equal(got, ref)
    p.close()
    o.close()


def test_recordings_and_servers_reject_caller_rdram(rdp_mi):
    stream = rs.config0_stream(num_triangles=8)
    p, host = _coherent(rdp_mi, stream)
    p.enqueue_commands(stream.frames[0].words)
    with pytest.raises(rdp_mi.RDPError):
        rdp_mi.Recording([p], batched=True)
    with pytest.raises(rdp_mi.RDPError):
        rdp_mi.Server([p])
    p.close()
