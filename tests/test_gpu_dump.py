"""RDPDUMP2 replay with per-draw validation, as the reference's
rdp-validate-dump does (rdp_validate_dump.cpp:84-198): the capture drives a
processor through the library's player (rdp_mi_dump_player_apply, the replay
driver of replayer_driver_parallel.cpp:124-205) and the oracle side by side;
after every draw call (or every SignalComplete in sync-only mode) TMEM, RDRAM
and hidden RDRAM must be identical, and every EndFrame's scanout (with the
replay driver's options: blend_previous_frame on, upscale_deinterlacing off)
must give the same image.  A capture taken from a processor must equal, byte
for byte, the capture the oracle writes for the same session."""
import os

import numpy as np
import pytest

import dumps
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


def _compare(proc, rep, what):
    proc.idle()
    gpu = (proc.read_rdram(), proc.read_hidden_rdram(), proc.read_tmem())
    total, lines = harness.compare(gpu, rep.state())
    assert total == 0, f"{what}: " + "; ".join(lines)


def _replay(rdp_mi, path, sync_only):
    import oracle
    checks = frames = 0
    with rdp_mi.DumpPlayer(path) as pl:
        proc = rdp_mi.CommandProcessor(pl.rdram_size)
        # The player's shadows start zeroed: so does the processor's state.
        proc.write_hidden_rdram(np.zeros(pl.rdram_size // 2, np.uint8))
        rep = dumps.OracleReplay(pl.rdram_size)
        opts_gpu = rdp_mi.ScanoutOptions.default(blend_previous_frame=1, upscale_deinterlacing=0)
        opts_ref = oracle.ScanoutOptions.default(blend_previous_frame=1, upscale_deinterlacing=0)
        draws = syncs = 0
        while True:
            ev = pl.next()
            if ev is None:
                break
            kind = ev.kind
            payload = None
            if kind == dumps.RDP_COMMAND:
                payload = (ev.op, pl.words())
            elif kind == dumps.SET_VI_REGISTER:
                payload = (ev.vi_register, ev.vi_value)
            pl.apply(proc)
            shadow_r = pl.rdram() if kind == dumps.UPDATE_DRAM_FLUSH else None
            shadow_h = pl.hidden_rdram() if kind == dumps.UPDATE_HIDDEN_DRAM_FLUSH else None
            rep.apply(kind, payload, shadow_r, shadow_h)
            if kind == dumps.END_FRAME:
                frames += 1
                img = proc.scanout_sync(opts_gpu)
                ref = rep.o.scanout(opts_ref)
                assert img.shape == ref.shape and np.array_equal(img, ref), f"scanout of frame {frames}"
            if kind == dumps.RDP_COMMAND and ev.op in dumps.DRAW_OPS:
                draws += 1
                if not sync_only:
                    _compare(proc, rep, f"frame {frames}, draw {draws}")
                    checks += 1
            if kind == dumps.SIGNAL_COMPLETE:
                syncs += 1
                if sync_only:
                    _compare(proc, rep, f"frame {frames}, sync {syncs}")
                    checks += 1
        _compare(proc, rep, "end of capture")
        rep.close()
        proc.close()
    return checks, frames


@pytest.mark.parametrize("name", dumps.GOLDEN_NAMES)
def test_replay_validates_every_draw(rdp_mi, name):
    path = dumps.golden_path(name)
    say(f"[case] replay {name}, per draw")
    try:
        checks, frames = _replay(rdp_mi, path, sync_only=name.startswith("c2"))
    finally:
        os.remove(path)
    say(f"[case] {name}: {checks} checks, {frames} frames")
    assert checks > 0
    if "vi" in name:
        assert frames == 2


@pytest.mark.parametrize("name", ["b3_two_cycle_rgba8888", "render_to_texture"])
def test_processor_capture_equals_oracle_capture(rdp_mi, tmp_path, name):
    stream = dumps.golden_streams()[name]
    say(f"[case] capture of {name} from a processor")
    path = str(tmp_path / "gpu.rdp")
    p = rdp_mi.CommandProcessor(rs.RDRAM_SIZE)
    p.write_rdram(stream.rdram_init if stream.rdram_init is not None else np.zeros(rs.RDRAM_SIZE // 4, np.uint32))
    p.write_hidden_rdram(np.full(rs.RDRAM_SIZE // 2, stream.hidden_init, np.uint8))
    p.begin_dump(path)
    for f in stream.frames:
        for cmd in dumps.commands(f.words):
            p.enqueue_command(cmd)
        if f.vi:
            for reg, val in f.vi.items():
                p.set_vi_register(reg, val)
            p.scanout_sync()
    p.idle()
    p.end_dump()
    p.close()
    with open(path, "rb") as fh:
        got = fh.read()
    ref_path = dumps.golden_path(name)
    try:
        with open(ref_path, "rb") as fh:
            assert got == fh.read()
    finally:
        os.remove(ref_path)
