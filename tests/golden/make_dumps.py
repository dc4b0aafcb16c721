"""Writes the RDPDUMP2 captures under tests/golden/dumps/ (gzip-compressed).

    python tests/golden/make_dumps.py

Each capture is one of the seeded command streams of
parallel-rdp_amd/rdp_streams.py, written the way the reference's capture hook
writes a session (tests/dumps.py, write_capture): the library's RDPDUMP2
writer, with the RDRAM state at each command-list start and scanout taken from
the oracle.  The files are fixtures for tests/test_dump.py and
tests/test_gpu_dump.py; regenerating them must give the same bytes.
"""
import gzip
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
for sub in ("oracle", "parallel-rdp_amd", "tests"):
    sys.path.insert(0, os.path.join(ROOT, sub))

import dumps  # noqa: E402


def main():
    os.makedirs(dumps.GOLDEN, exist_ok=True)
    for name, stream in dumps.golden_streams().items():
        raw = os.path.join(dumps.GOLDEN, name + ".rdp")
        dumps.write_capture(stream, raw)
        with open(raw, "rb") as fh:
            data = fh.read()
        os.remove(raw)
        # mtime 0 keeps the compressed bytes reproducible.
        with open(raw + ".gz", "wb") as fh, gzip.GzipFile(fileobj=fh, mode="wb", mtime=0, compresslevel=9) as gz:
            gz.write(data)
        print(f"{name}: {len(data)} bytes, {os.path.getsize(raw + '.gz')} compressed")


if __name__ == "__main__":
    main()
