"""Pins the lookup tables generated from closed forms (oracle and library)
against SHA-256 digests of the tables the reference carries as data
(tests/golden/reference_tables.json, written by tests/golden/make_golden.py)."""
import hashlib
import json
import os

import numpy as np

GOLDEN = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_tables.json")))


def test_blender_divider_matches_reference(oracle_mod):
    t = oracle_mod.blender_divider_table()
    assert t.size == GOLDEN["blender_divider"]["bytes"]
    assert hashlib.sha256(t.tobytes()).hexdigest() == GOLDEN["blender_divider"]["sha256"]


def test_blender_divider_is_exact_division_in_range(oracle_mod):
    # Independent of the reference data: for divisors 1..8 and quotients
    # below 256 the hardware divider must agree with integer division.
    t = oracle_mod.blender_divider_table()
    for d in range(1, 9):
        for n in range(0, min(2048, 256 * d)):
            assert t[(d << 11) | n] == n // d, (d, n)


def test_gamma_matches_reference(oracle_mod):
    t = oracle_mod.gamma_table()
    assert hashlib.sha256(t.tobytes()).hexdigest() == GOLDEN["gamma"]["sha256"]
    # closed form spot checks: sqrt curve truncated to even
    assert t[0] == 0 and t[1] == 16 and t[4] == 32 and t[255] == 254


def test_perspective_table_matches_reference(oracle_mod):
    t = oracle_mod.perspective_table()
    assert t.shape == (64, 2)
    assert hashlib.sha256(t.astype("<i2").tobytes()).hexdigest() == GOLDEN["perspective"]["sha256"]
    # the table is 2^20 / (64 + i): first entry is exactly 1.0 in 1.14
    assert t[0, 0] == 0x4000


def test_dither_matrices_match_reference():
    magic = [0, 6, 1, 7, 4, 2, 5, 3, 3, 5, 2, 4, 7, 1, 6, 0]
    bayer = [0, 4, 1, 5, 4, 0, 5, 1, 3, 7, 2, 6, 7, 3, 6, 2]
    b = bytes(magic + bayer)
    assert hashlib.sha256(b).hexdigest() == GOLDEN["dither"]["sha256"]
    # Both are permutations of 0..7 twice; Bayer quadrants sum per 2x2 block.
    assert sorted(magic) == sorted(list(range(8)) * 2)
    assert sorted(bayer) == sorted(list(range(8)) * 2)
