"""Known-answer tests of the device's counter-based generators.

Philox-4x32-10 (source and interaction sampling) and Philox-2x32-10 (the
Woodcock steps of the queue kernel) against the published known-answer
vectors of Salmon et al., "Parallel random numbers: as easy as 1, 2, 3"
(SC'11; the Random123 distribution's kat_vectors file).  The reference draws
from std::mt19937 (probability_dist.cpp:4), so no reference stream exists to
compare with; these vectors fix the generator itself.  The CPU test checks the
vectors against a plain restatement of the published round function (the
restatement and the vectors must agree before the device is compared with
either); the GPU test runs the device's own round functions through the
C-ABI."""
import numpy as np
import pytest

import midsx_amd as M

MASK = 0xFFFFFFFF

# (counter, key, output) from the published known-answer vectors
KAT4 = [
    ((0, 0, 0, 0), (0, 0), (0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8)),
    ((MASK,) * 4, (MASK, MASK), (0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD)),
    ((0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344), (0xA4093822, 0x299F31D0),
     (0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1)),
]
KAT2 = [
    ((0, 0), 0, (0xFF1DAE59, 0x6CD10DF2)),
    ((MASK, MASK), MASK, (0x2C3F628B, 0xAB4FD7AD)),
    ((0x243F6A88, 0x85A308D3), 0x13198A2E, (0xDD7CE038, 0xF62A4C12)),
]


def _mulhilo(a, b):
    p = a * b
    return p >> 32, p & MASK


def philox4(ctr, key, rounds=10):
    x0, x1, x2, x3 = ctr
    k0, k1 = key
    for _ in range(rounds):
        h0, l0 = _mulhilo(0xD2511F53, x0)
        h1, l1 = _mulhilo(0xCD9E8D57, x2)
        x0, x1, x2, x3 = h1 ^ x1 ^ k0, l1, h0 ^ x3 ^ k1, l0
        k0, k1 = (k0 + 0x9E3779B9) & MASK, (k1 + 0xBB67AE85) & MASK
    return (x0, x1, x2, x3)


def philox2(ctr, key, rounds=10):
    x0, x1 = ctr
    k = key
    for _ in range(rounds):
        h, l = _mulhilo(0xD256D193, x0)
        x0, x1 = h ^ k ^ x1, l
        k = (k + 0x9E3779B9) & MASK
    return (x0, x1)


def test_restatement_reproduces_published_vectors():
    for c, k, out in KAT4:
        assert philox4(c, k) == out
    for c, k, out in KAT2:
        assert philox2(c, k) == out
    # one round fewer gives other words: the vectors pin the round count
    assert philox4(KAT4[2][0], KAT4[2][1], 9) != KAT4[2][2]


@pytest.mark.gpu
def test_device_philox_known_answers():
    ctx = M.Context(0)
    c4 = np.array([c for c, _, _ in KAT4], np.uint32)
    k4 = np.array([k for _, k, _ in KAT4], np.uint32)
    got4 = ctx.philox(c4, k4, words=4)
    assert [tuple(int(v) for v in r) for r in got4] == [o for _, _, o in KAT4]
    c2 = np.array([c for c, _, _ in KAT2], np.uint32)
    k2 = np.array([k for _, k, _ in KAT2], np.uint32)
    got2 = ctx.philox(c2, k2, words=2)
    assert [tuple(int(v) for v in r) for r in got2] == [o for _, _, o in KAT2]
    # and on random counters and keys, word for word against the restatement
    rng = np.random.default_rng(8)
    c = rng.integers(0, 1 << 32, size=(4096, 4), dtype=np.uint64).astype(np.uint32)
    k = rng.integers(0, 1 << 32, size=(4096, 2), dtype=np.uint64).astype(np.uint32)
    got = ctx.philox(c, k, words=4)
    for i in range(0, 4096, 97):
        assert tuple(int(v) for v in got[i]) == philox4(tuple(int(v) for v in c[i]), tuple(int(v) for v in k[i]))
    got = ctx.philox(c[:, :2], k[:, 0], words=2)
    for i in range(0, 4096, 97):
        assert tuple(int(v) for v in got[i]) == philox2(tuple(int(v) for v in c[i, :2]), int(k[i, 0]))
    ctx.close()
