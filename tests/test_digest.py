"""The state digest that checks served frames one by one: the oracle's C
digest equals a numpy restatement of the documented definition, and any
single changed byte in RDRAM, hidden RDRAM or TMEM changes it."""
import numpy as np

M64 = (1 << 64) - 1
KEYS = (0x243F6A8885A308D3, 0x13198A2E03707344, 0xA4093822299F31D0)


def _mix64(z):
    z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
    z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    return z ^ (z >> np.uint64(31))


def _region(words, key):
    i = np.arange(words.size, dtype=np.uint64)
    with np.errstate(over="ignore"):
        terms = _mix64(np.uint64(key) ^ ((i << np.uint64(32)) | words.astype(np.uint64)))
    return int(terms.sum(dtype=np.uint64))


def digest(rdram, hidden, tmem):
    return (_region(rdram.view(np.uint32), KEYS[0]) + _region(hidden.view(np.uint32), KEYS[1]) +
            _region(tmem.view(np.uint32), KEYS[2])) & M64


def test_oracle_digest_matches_definition(oracle_mod):
    o = oracle_mod.Oracle(1 << 20)
    rng = np.random.default_rng(4)
    o.rdram[:] = rng.integers(0, 1 << 32, size=o.rdram.size, dtype=np.uint64).astype(np.uint32)
    o.hidden[:] = rng.integers(0, 4, size=o.hidden.size).astype(np.uint8)
    o.tmem[:] = rng.integers(0, 1 << 16, size=2048).astype(np.uint16)
    want = digest(o.rdram, o.hidden, o.tmem)
    assert o.digest() == want
    for arr, idx in ((o.rdram.view(np.uint8), 12345), (o.hidden, 777), (o.tmem.view(np.uint8), 4000)):
        old = int(arr[idx])
        arr[idx] = old ^ 1
        assert o.digest() != want
        arr[idx] = old
    assert o.digest() == want
    o.close()
