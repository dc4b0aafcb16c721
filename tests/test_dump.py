"""RDPDUMP2 captures on the CPU: the committed captures parse with an
independent reader of the format (rdp_dump.cpp:43-200), regenerate to the same
bytes, read the same through the library's player, and replay on the oracle
to the state of the streams they were written from.  The library's writer
sends only changed 4 KiB blocks, and its player rejects malformed files."""
import gzip
import os
import struct

import numpy as np
import pytest

import dumps
import harness
import rdp_streams as rs


@pytest.fixture(scope="module")
def rdp_mi():
    import rdp_mi as mod
    mod.load_library()
    return mod


def _golden_bytes(name):
    with gzip.open(os.path.join(dumps.GOLDEN, name + ".rdp.gz"), "rb") as fh:
        return fh.read()


@pytest.mark.parametrize("name", dumps.GOLDEN_NAMES)
def test_golden_capture_parses(rdp_mi, name):
    rdram_size, hidden_size, recs = dumps.read_records(_golden_bytes(name))
    assert rdram_size == rs.RDRAM_SIZE and hidden_size == rs.RDRAM_SIZE // 2
    assert recs[-1][0] == dumps.EOF and all(k != dumps.EOF for k, _ in recs[:-1])
    n_cmd = n_sync = 0
    for kind, payload in recs:
        if kind == dumps.RDP_COMMAND:
            op, words = payload
            assert op == (int(words[0]) >> 24) & 63 and op != dumps.SYNC_FULL
            assert words.size == rdp_mi.COMMAND_LENGTHS[op]
            n_cmd += 1
        elif kind in (dumps.UPDATE_DRAM, dumps.UPDATE_HIDDEN_DRAM):
            off, blk = payload
            assert off % 4096 == 0 and len(blk) == 4096
        elif kind == dumps.SIGNAL_COMPLETE:
            n_sync += 1
    assert n_cmd > 0 and n_sync > 0


@pytest.mark.parametrize("name", ["c0_shaded_64", "b3_two_cycle_rgba8888", "render_to_texture"])
def test_capture_regenerates_to_the_committed_bytes(rdp_mi, tmp_path, name):
    path = str(tmp_path / "cap.rdp")
    dumps.write_capture(dumps.golden_streams()[name], path)
    with open(path, "rb") as fh:
        assert fh.read() == _golden_bytes(name)


@pytest.mark.parametrize("name", dumps.GOLDEN_NAMES)
def test_library_player_reads_what_the_reader_reads(rdp_mi, name):
    data = _golden_bytes(name)
    _, _, recs = dumps.read_records(data)
    path = dumps.golden_path(name)
    try:
        with rdp_mi.DumpPlayer(path) as pl:
            assert pl.rdram_size == rs.RDRAM_SIZE and pl.hidden_rdram_size == rs.RDRAM_SIZE // 2
            shadow = [np.zeros(rs.RDRAM_SIZE, np.uint8), np.zeros(rs.RDRAM_SIZE // 2, np.uint8)]
            for kind, payload in recs:
                ev = pl.next()
                if kind == dumps.EOF:
                    assert ev is None
                    break
                assert ev.kind == kind
                if kind == dumps.RDP_COMMAND:
                    assert ev.op == payload[0] and np.array_equal(pl.words(), payload[1])
                elif kind == dumps.SET_VI_REGISTER:
                    assert (ev.vi_register, ev.vi_value) == payload
                elif kind in (dumps.UPDATE_DRAM, dumps.UPDATE_HIDDEN_DRAM):
                    off, blk = payload
                    assert (ev.offset, ev.size) == (off, len(blk))
                    shadow[kind == dumps.UPDATE_HIDDEN_DRAM][off:off + len(blk)] = np.frombuffer(blk, np.uint8)
                elif kind == dumps.UPDATE_DRAM_FLUSH:
                    assert np.array_equal(pl.rdram().view(np.uint8), shadow[0])
                elif kind == dumps.UPDATE_HIDDEN_DRAM_FLUSH:
                    assert np.array_equal(pl.hidden_rdram(), shadow[1])
            # Rewinding starts over with cleared shadows.
            pl.rewind()
            assert not pl.rdram().any()
            first = pl.next()
            assert first is not None and first.kind == recs[0][0]
    finally:
        os.remove(path)


@pytest.mark.parametrize("name", dumps.GOLDEN_NAMES)
def test_capture_replays_on_the_oracle_to_the_stream_state(rdp_mi, name):
    """Replaying a capture (blocks, commands, SignalComplete as a flush) gives
    the state of rendering its stream directly."""
    stream = dumps.golden_streams()[name]
    rdram_size, _, recs = dumps.read_records(_golden_bytes(name))
    rep = dumps.OracleReplay(rdram_size)
    shadow = [np.zeros(rdram_size, np.uint8), np.zeros(rdram_size // 2, np.uint8)]
    for kind, payload in recs:
        if kind in (dumps.UPDATE_DRAM, dumps.UPDATE_HIDDEN_DRAM):
            off, blk = payload
            shadow[kind == dumps.UPDATE_HIDDEN_DRAM][off:off + len(blk)] = np.frombuffer(blk, np.uint8)
        rep.apply(kind, payload, shadow[0].view(np.uint32), shadow[1])
    got = rep.state()
    rep.close()
    ref = harness.run_oracle(stream)
    total, lines = harness.compare(got, ref)
    assert total == 0, "; ".join(lines)


def test_writer_sends_only_changed_blocks(rdp_mi, tmp_path):
    path = str(tmp_path / "w.rdp")
    a = np.zeros(rs.RDRAM_SIZE, np.uint8)
    a[5000] = 7
    a[3 * 4096 + 1] = 9
    h = np.zeros(rs.RDRAM_SIZE // 2, np.uint8)
    with rdp_mi.DumpWriter(path, rs.RDRAM_SIZE, rs.RDRAM_SIZE // 2) as w:
        w.flush_rdram(a)
        w.flush_rdram(a)  # nothing changed: only the flush record
        a[5001] = 1
        w.flush_rdram(a)
        w.flush_hidden_rdram(h)
        w.set_vi_register(3, 0x12345678)
        w.command(np.array([0x29000000, 0], np.uint32))
        w.end_frame()
    with open(path, "rb") as fh:
        _, _, recs = dumps.read_records(fh.read())
    kinds = [k for k, _ in recs]
    assert kinds == [dumps.UPDATE_DRAM, dumps.UPDATE_DRAM, dumps.UPDATE_DRAM_FLUSH, dumps.UPDATE_DRAM_FLUSH,
                     dumps.UPDATE_DRAM, dumps.UPDATE_DRAM_FLUSH, dumps.UPDATE_HIDDEN_DRAM_FLUSH,
                     dumps.SET_VI_REGISTER, dumps.RDP_COMMAND, dumps.END_FRAME, dumps.EOF]
    assert [recs[i][1][0] for i in (0, 1, 4)] == [4096, 3 * 4096, 4096]
    assert recs[7][1] == (3, 0x12345678)


def test_player_rejects_malformed_captures(rdp_mi, tmp_path):
    def write(name, data):
        p = str(tmp_path / name)
        with open(p, "wb") as fh:
            fh.write(data)
        return p

    head = b"RDPDUMP2" + struct.pack("<II", rs.RDRAM_SIZE, rs.RDRAM_SIZE // 2)
    with pytest.raises(rdp_mi.RDPError):
        rdp_mi.DumpPlayer(write("magic.rdp", b"RDPDUMP1" + head[8:]))
    with pytest.raises(rdp_mi.RDPError):
        rdp_mi.DumpPlayer(write("size.rdp", b"RDPDUMP2" + struct.pack("<II", 3 << 20, 4 << 20)))
    # A block past the end of RDRAM, a truncated command and an unknown kind.
    for name, rec in (("block", struct.pack("<III", 1, rs.RDRAM_SIZE - 4, 8) + b"\0" * 8),
                      ("cmd", struct.pack("<IIII", 2, 0x29, 2, 0)),
                      ("kind", struct.pack("<I", 77))):
        with rdp_mi.DumpPlayer(write(name + ".rdp", head + rec)) as pl:
            with pytest.raises(rdp_mi.RDPError):
                pl.next()
    # A file that ends on a record boundary ends the capture.
    with rdp_mi.DumpPlayer(write("end.rdp", head + struct.pack("<I", 5))) as pl:
        assert pl.next().kind == dumps.SIGNAL_COMPLETE
        assert pl.next() is None
