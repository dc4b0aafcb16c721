"""TMEM uploads on the GPU against the oracle (LoadTile, LoadBlock, LoadTLUT
with the ramp pattern of the reference's load conformance test)."""
import numpy as np
import pytest

import rdp_streams as rs
import test_oracle_tmem as t

pytestmark = pytest.mark.gpu


def gpu_load(words):
    import rdp_mi
    p = rdp_mi.CommandProcessor(rs.RDRAM_SIZE)
    p.write_rdram(t.ramp_rdram())
    p.enqueue_commands(np.asarray(words, dtype=np.uint32))
    p.idle()
    out = p.read_tmem()
    p.close()
    return out


CASES = []
for size in (rs.SIZ_4, rs.SIZ_8, rs.SIZ_16, rs.SIZ_32):
    for width, height, offset, stride in ((32, 7, 0, 128), (16, 3, 0x800, 256), (12, 5, 64, 64), (64, 40, 0, 136),
                                          (100, 50, 8, 200), (3, 1, 0, 8)):
        CASES.append(("tile", size, width, height, offset, stride))
for pixels, dxt in ((64, 0), (256, 0x800), (512, 0x400), (2048, 0x200), (100, 0x155), (1000, 0x7ff)):
    for size in (rs.SIZ_8, rs.SIZ_16, rs.SIZ_32):
        CASES.append(("block", size, pixels, dxt, 0, 0))
for entries, offset in ((16, 0x800), (256, 0x800), (16, 0xC00), (7, 0x800), (256, 0)):
    CASES.append(("tlut", rs.SIZ_16, entries, offset, 0, 0))


def words_for(case):
    kind, size, a, b, offset, stride = case
    cl = rs.CommandList()
    if kind == "tile":
        return t.cmds_tile(a, b, offset, stride, size, 128)
    if kind == "block":
        cl.set_texture_image(t.TEX_BASE, 256, size, rs.FMT_RGBA)
        cl.set_tile(0, rs.FMT_RGBA, size, 0x100, 0)
        cl.load_block(0, 0, 0, a - 1, b)
        return cl.words()
    cl.set_texture_image(t.TEX_BASE, 256, rs.SIZ_16, rs.FMT_RGBA)
    cl.set_tile(7, rs.FMT_RGBA, rs.SIZ_4, b, 0)
    cl.load_tlut(7, 0, 0, (a - 1) << 2, 0)
    return cl.words()


@pytest.mark.parametrize("case", CASES, ids=[f"{c[0]}-{c[1]}-{c[2]}-{c[3]}-{c[4]}-{c[5]}" for c in CASES])
def test_tmem_upload_matches_oracle(oracle_mod, case):
    print(f"[case] tmem {case}", flush=True)
    words = words_for(case)
    ref, _ = t.run_load(oracle_mod, words)
    got = gpu_load(words)
    diff = np.nonzero(got != ref)[0]
    assert diff.size == 0, f"{diff.size} TMEM halfwords differ, first at {int(diff[0])}"
