"""Builds librdp_mi.so (host front end + gfx950 kernels) in-tree.

    python parallel-rdp_amd/build.py

Compiles with hipcc for gfx950 only.  The shared library lands next to this
file so it travels with the source tree.
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "librdp_mi.so")
BUILD_DIR = os.path.join(HERE, "build")

SOURCES = ["rdp_frontend.cpp", "rdp_dump.cpp", "rdp_mi.cpp", "rdp_kernels.hip", "rdp_vi.hip", "rdp_digest.hip", "rdp_view.hip"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

COMMON = [
    "-O3", "-fPIC", "-std=c++17", "-fwrapv", "-Wall", "-Wno-unused-function",
    "--offload-arch=gfx950",
    # Keep the compiler from forming packed FP32 instructions.
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def _run(cmd):
    print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build(verbose: bool = True) -> str:
    os.makedirs(BUILD_DIR, exist_ok=True)
    objs = []
    for src in SOURCES:
        path = os.path.join(CSRC, src)
        if not os.path.exists(path):
            continue
        obj = os.path.join(BUILD_DIR, os.path.splitext(src)[0] + ".o")
        cmd = [HIPCC] + COMMON + ["-c", path, "-o", obj]
        if src.endswith(".cpp"):
            cmd = [HIPCC] + COMMON + ["-x", "hip", "-c", path, "-o", obj]
        _run(cmd)
        objs.append(obj)
    tmp = OUT + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs)
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    build()
    sys.exit(0)
