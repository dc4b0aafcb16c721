"""Writes tests/golden/reference_tables.json: SHA-256 digests of the data
tables the reference carries in its sources, so the tables this project
generates from closed forms can be pinned without copying them.

Reads (as text) parallel-rdp/luts.hpp (blender_lut, gamma_table),
parallel-rdp/shaders/perspective.h (perspective_table) and
parallel-rdp/shaders/dither.h (dither_matrices) under /root/reference.
Run in the build container:  python tests/golden/make_golden.py
"""
import hashlib
import json
import os
import re
import struct

REF = "/root/reference/parallel-rdp"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "reference_tables.json")


def _numbers(text):
    return [int(x) for x in re.findall(r"-?\d+", text)]


def main():
    luts = open(os.path.join(REF, "luts.hpp")).read()
    m = re.search(r"blender_lut\[[^\]]*\]\s*=\s*\{(.*?)\};", luts, re.S)
    blender = bytes(_numbers(m.group(1)))
    m = re.search(r"gamma_table\[[^\]]*\]\s*=\s*\{(.*?)\};", luts, re.S)
    gamma = bytes(_numbers(m.group(1)))

    persp_src = open(os.path.join(REF, "shaders", "perspective.h")).read()
    m = re.search(r"perspective_table\[64\]\s*=\s*i16x2\[\]\((.*?)\);", persp_src, re.S)
    pairs = re.findall(r"i16x2\((0x[0-9a-fA-F]+),\s*(-?\d+)\s*\*\s*(\d+)\)", m.group(1))
    persp = b"".join(struct.pack("<hh", int(a, 16), int(b) * int(c)) for a, b, c in pairs)

    dither_src = open(os.path.join(REF, "shaders", "dither.h")).read()
    m = re.search(r"dither_matrices\[2\]\[16\]\s*=\s*u8\[\]\[\]\((.*?)\);", dither_src, re.S)
    dither = bytes(int(x) for x in re.findall(r"U8_C\((\d+)\)", m.group(1)))

    out = {
        "source": "Themaister/parallel-rdp (commit unknown; /root/reference has no .git)",
        "blender_divider": {"bytes": len(blender), "sha256": hashlib.sha256(blender).hexdigest(),
                            "from": "parallel-rdp/luts.hpp blender_lut"},
        "gamma": {"bytes": len(gamma), "sha256": hashlib.sha256(gamma).hexdigest(),
                  "from": "parallel-rdp/luts.hpp gamma_table"},
        "perspective": {"entries": len(pairs), "sha256": hashlib.sha256(persp).hexdigest(),
                        "layout": "64 x (int16 reciprocal, int16 slope) little-endian",
                        "from": "parallel-rdp/shaders/perspective.h perspective_table"},
        "dither": {"bytes": len(dither), "sha256": hashlib.sha256(dither).hexdigest(),
                   "layout": "2 x 16 uint8 (magic square, Bayer)",
                   "from": "parallel-rdp/shaders/dither.h dither_matrices"},
    }
    with open(OUT, "w") as f:
        json.dump(out, f, indent=2)
        f.write("\n")
    print(json.dumps(out, indent=2))


if __name__ == "__main__":
    main()
