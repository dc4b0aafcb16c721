"""Host front end (no device): its reference render-pass boundaries agree with
the oracle's, and it only leaves a pass open across such a boundary when the
boundary is unobservable (never for I8 colour images)."""
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
    out = os.path.join(tempfile.mkdtemp(prefix="fe_probe_"), "libfe_probe.so")
    subprocess.run(["g++", "-std=c++17", "-O2", "-fPIC", "-shared", "-fwrapv", "-o", out,
                    os.path.join(ROOT, "tests", "frontend_probe.cpp"),
                    os.path.join(ROOT, "parallel-rdp_amd", "csrc", "rdp_frontend.cpp")], check=True)
    lib = ctypes.CDLL(out)
    vp, sz = ctypes.c_void_p, ctypes.c_size_t
    lib.fp_create.restype = vp
    lib.fp_create.argtypes = [ctypes.c_uint32]
    lib.fp_destroy.argtypes = [vp]
    lib.fp_process.argtypes = [vp, vp, sz]
    lib.fp_flush.argtypes = [vp]
    lib.fp_pass_ends.restype = sz
    lib.fp_pass_ends.argtypes = [vp, vp, sz]
    lib.fp_passes.restype = sz
    lib.fp_passes.argtypes = [vp, vp, vp, vp, sz]
    return lib


def frontend_run(lib, stream):
    h = lib.fp_create(rs.RDRAM_SIZE)
    for f in stream.frames:
        w = np.ascontiguousarray(f.words, dtype=np.uint32)
        lib.fp_process(h, w.ctypes.data, w.size)
    lib.fp_flush(h)
    n = lib.fp_pass_ends(h, None, 0)
    ends = np.zeros(n, dtype=np.uint64)
    lib.fp_pass_ends(h, ends.ctypes.data, n)
    m = lib.fp_passes(h, None, None, None, 0)
    first, last, ups = (np.zeros(m, dtype=np.uint32) for _ in range(3))
    lib.fp_passes(h, first.ctypes.data, last.ctypes.data, ups.ctypes.data, m)
    lib.fp_destroy(h)
    return ends, first, last


def oracle_ends(oracle_mod, stream):
    o = oracle_mod.Oracle(rs.RDRAM_SIZE)
    if stream.rdram_init is not None:
        o.rdram[:] = stream.rdram_init
    for f in stream.frames:
        o.process(f.words)
    o.flush()
    ends = o.pass_ends()
    o.close()
    return ends


STREAMS = ([("fuzz", k) for k in range(48)] + [("batch", i) for i in (0, 1, 2, 5, 7, 12)] +
           [("c0", 0), ("c1", 0)])


def make(kind, k):
    if kind == "fuzz":
        return rs.fuzz_stream(k, segments=24)
    if kind == "batch":
        return rs.batch_stream(k)
    if kind == "c0":
        return rs.config0_stream(num_triangles=1500)
    return rs.config1_stream(frames=1, tris_per_frame=1500)


@pytest.mark.parametrize("kind,k", STREAMS)
def test_pass_boundaries_match_oracle(probe, oracle_mod, kind, k):
    s = make(kind, k)
    ends, first, last = frontend_run(probe, s)
    ref = oracle_ends(oracle_mod, s)
    assert np.array_equal(ends, ref), f"{s.name}: front end {ends[:20]} vs oracle {ref[:20]}"
    # Device passes are unions of whole reference passes: between two device
    # passes there is a reference boundary (draw e-1 ends a pass, draw e
    # starts one, in primitive-counter terms).
    ref_list = [int(x) for x in ref]
    spans = [(int(a), int(b)) for a, b in zip(first, last) if a != 0xFFFFFFFF]
    for (a0, b0), (a1, b1) in zip(spans, spans[1:]):
        assert any(b0 < e <= a1 for e in ref_list), f"{s.name}: device split between {b0} and {a1}"
    if s.color_size == rs.SIZ_8:
        # I8: no reference boundary falls inside a device pass.
        for a, b in spans:
            assert not any(a < e <= b for e in ref_list), f"{s.name}: pass {a}..{b} spans a boundary"


def test_large_stream_splits_at_256_primitives(probe, oracle_mod):
    s = rs.config0_stream(num_triangles=1000)
    ends, first, last = frontend_run(probe, s)
    assert len(ends) >= 4
    assert np.array_equal(ends, oracle_ends(oracle_mod, s))
