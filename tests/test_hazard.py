"""The bound that lets a round's TMEM replay start before the previous round's
raster (parallel-rdp_amd/csrc/rdp_hazard.hpp, upload_read_range) must contain
every RDRAM byte the TMEM upload plan reads (rdp_tmem_plan.hpp, the planning
half of the gather kernel), or a replay could read memory a raster still
writes.  Checked on the host over random-state streams (every load kind,
format and size drawn at random), the batch workloads and render-to-texture
streams."""
import ctypes
import os
import subprocess
import tempfile

import numpy as np
import pytest

import rdp_streams as rs

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def probe():
    tmp = tempfile.mkdtemp(prefix="hz_probe_")
    out = os.path.join(tmp, "libhz_probe.so")
    objs = []
    for src in ("rdp_oracle.c", "vi_oracle.c"):
        obj = os.path.join(tmp, src + ".o")
        subprocess.run(["gcc", "-std=c99", "-O2", "-fPIC", "-fwrapv", "-c", "-o", obj,
                        os.path.join(ROOT, "oracle", src)], check=True)
        objs.append(obj)
    subprocess.run(["g++", "-std=c++17", "-O2", "-fPIC", "-shared", "-fwrapv", "-o", out,
                    os.path.join(ROOT, "tests", "hazard_probe.cpp"),
                    os.path.join(ROOT, "parallel-rdp_amd", "csrc", "rdp_frontend.cpp")] + objs + ["-lm"], check=True)
    lib = ctypes.CDLL(out)
    lib.hz_check.restype = ctypes.c_int
    lib.hz_check.argtypes = [ctypes.c_void_p, ctypes.c_size_t, ctypes.c_uint32, ctypes.c_void_p]
    lib.hz_write_check.restype = ctypes.c_int
    lib.hz_write_check.argtypes = [ctypes.c_void_p, ctypes.c_size_t, ctypes.c_uint32, ctypes.c_void_p, ctypes.c_uint32,
                                   ctypes.c_void_p]
    return lib


def check(lib, stream):
    w = np.ascontiguousarray(stream.all_words(), dtype=np.uint32)
    out = np.zeros(4, dtype=np.uint64)
    assert lib.hz_check(w.ctypes.data, w.size, rs.RDRAM_SIZE, out.ctypes.data) == 0
    return [int(v) for v in out]


@pytest.mark.parametrize("seed", range(0, 40))
def test_upload_reads_stay_in_bound_fuzz(probe, seed):
    uploads, read, outside, refused = check(probe, rs.fuzz_stream(0x4A2D000 + seed, segments=24))
    assert outside == 0, f"{outside} of {read} bytes read outside the bound"


def test_upload_reads_stay_in_bound_workloads(probe):
    total_uploads = total_read = 0
    streams = [rs.batch_stream(i, max_tris=600) for i in range(0, 64, 3)] + [rs.config1_stream(frames=1),
                                                                             rs.config2_stream(frames=1)]
    for s in streams:
        uploads, read, outside, refused = check(probe, s)
        assert outside == 0, f"{s.name}: {outside} of {read} bytes read outside the bound"
        assert refused == 0, f"{s.name}: {refused} uploads without a bound"
        total_uploads += uploads
        total_read += read
    # The workloads load textures and TLUTs: the check saw real reads.
    assert total_uploads > 100 and total_read > 100000


def test_upload_reads_stay_in_bound_render_to_texture(probe):
    import test_gpu_rounds as rounds
    for i in range(6):
        uploads, read, outside, refused = check(probe, rounds.texture_stream(100 + i, i % 2 == 0))
        assert uploads >= 1 and read > 0
        assert outside == 0


def write_check(lib, stream):
    w = np.ascontiguousarray(stream.all_words(), dtype=np.uint32)
    init = None if stream.rdram_init is None else np.ascontiguousarray(stream.rdram_init, dtype=np.uint32)
    out = np.zeros(3, dtype=np.uint64)
    assert lib.hz_write_check(w.ctypes.data, w.size, rs.RDRAM_SIZE, None if init is None else init.ctypes.data,
                              stream.hidden_init, out.ctypes.data) == 0
    return [int(v) for v in out]


# pass_write_ranges (rdp_hazard.hpp) is the other half of the early-replay
# check: a raster must write nothing outside it.  The oracle renders each
# stream and every byte a device pass changes must fall inside that pass's
# colour and depth ranges.
@pytest.mark.parametrize("seed", range(0, 24))
def test_raster_writes_stay_in_bound_fuzz(probe, seed):
    passes, changed, outside = write_check(probe, rs.fuzz_stream(0x51A7000 + seed, segments=16))
    assert outside == 0, f"{outside} of {changed} changed bytes outside the pass's write ranges ({passes} passes)"


def test_raster_writes_stay_in_bound_workloads(probe):
    streams = [rs.batch_stream(i, max_tris=400) for i in range(0, 64, 5)] + [rs.config0_stream(num_triangles=64)]
    total = 0
    for s in streams:
        passes, changed, outside = write_check(probe, s)
        assert outside == 0, f"{s.name}: {outside} of {changed} changed bytes outside the write ranges"
        total += changed
    assert total > 100000


def test_raster_writes_stay_in_bound_alias_and_render_to_texture(probe):
    import test_gpu_rounds as rounds
    streams = [rounds.texture_stream(300 + i, i % 2 == 0) for i in range(4)]
    streams += [rs.alias_stream(0x5300 + 4 * fmt + 2 * cycle + partial, fmt, cycle, partial)
                for fmt in (rs.FMT_RGBA, rs.FMT_IA) for cycle in (0, 1) for partial in (False, True)]
    for s in streams:
        passes, changed, outside = write_check(probe, s)
        assert passes >= 1 and changed > 0, s.name
        assert outside == 0, f"{s.name}: {outside} of {changed} changed bytes outside the write ranges"
