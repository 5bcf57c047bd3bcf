"""CPU check of the rule behind the pair-once Oseen kernel's 32-bit
separations (csrc/direct.hip, boundary pairs): with floor high words
h = q >> 32 of 64-bit ring positions q, the 32-bit difference
c = h_j - h_i (mod 2^32), read as int32, equals hi = floor(D / 2^32) of the
64-bit difference D = q_j - q_i (read as int64) or hi + 1 (the low words'
borrow) -- one unit of the separation -- except when c == 2^31 and
hi == 2^31 - 1: the int32 wrap, where the 32-bit value names the other
periodic image.  So a pair can change image only if c == 2^31, the pairs the
kernel's lookup finds.  Brute force on random words and around half the
ring."""

from __future__ import annotations

import numpy as np

M64 = (1 << 64) - 1


def _as_i32(c: int) -> int:
    return c - (1 << 32) if c >= (1 << 31) else c


def _as_i64(d: int) -> int:
    return d - (1 << 64) if d >= (1 << 63) else d


def _classify(qi: int, qj: int) -> str:
    c = ((qj >> 32) - (qi >> 32)) & 0xFFFFFFFF
    hi = _as_i64((qj - qi) & M64) >> 32  # floor of D / 2^32
    c32 = _as_i32(c)
    if c32 - hi in (0, 1):
        return "boundary-agree" if c == 1 << 31 else "agree"
    assert c == 1 << 31 and hi == (1 << 31) - 1, (qi, qj)
    return "wrap"


def test_random_words():
    rng = np.random.default_rng(11)
    q = rng.integers(0, 1 << 63, size=(20000, 2), dtype=np.uint64) * 2 + rng.integers(
        0, 2, size=(20000, 2), dtype=np.uint64)
    for a, b in q.tolist():
        _classify(int(a), int(b))


def test_pairs_around_half_the_ring_and_zero():
    """Partners at q_i + 2^63 + delta and q_i + delta for small delta: the only
    departures are wraps with c == 2^31, and both wraps and non-wrapping
    c == 2^31 pairs occur."""
    rng = np.random.default_rng(12)
    seen = {"agree": 0, "boundary-agree": 0, "wrap": 0}
    deltas = (-(1 << 33), -(1 << 32) - 1, -(1 << 32), -5, -1, 0, 1, 5, 1 << 32, (1 << 32) + 1,
              1 << 33)
    for qi in rng.integers(0, 1 << 63, size=400, dtype=np.uint64).tolist():
        qi = int(qi) * 2 + 1
        for base in (1 << 63, 0):
            for delta in deltas:
                seen[_classify(qi, (qi + base + delta) & M64)] += 1
    assert seen["wrap"] > 0 and seen["boundary-agree"] > 0 and seen["agree"] > 0
