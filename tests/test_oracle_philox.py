"""Pins the oracle's Philox stream (CPU only).

Known-answer vectors are the Random123 ones the reference's tests hold
(tests/test_philox.py:7-18 of the reference); the stream-layout facts are those of
tests/test_philox.py:29-34,45-49; the C restatement must agree with the NumPy one.
"""

import numpy as np
import pytest

from oracle import c_oracle as CO
from oracle import ising_oracle as O

KAT = [
    ((0, 0, 0, 0), (0, 0), (0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8)),
    ((0xFFFFFFFF,) * 4, (0xFFFFFFFF,) * 2, (0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD)),
    ((0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344), (0xA4093822, 0x299F31D0),
     (0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1)),
]


@pytest.mark.parametrize("ctr,key,expected", KAT)
def test_known_answers(ctr, key, expected):
    assert tuple(int(x) for x in O.philox4x32(*ctr, *key)) == expected


@pytest.mark.parametrize("ctr,key,expected", KAT)
def test_c_restatement_known_answers(ctr, key, expected):
    out = CO.philox(np.array(ctr), np.array(key))
    assert tuple(int(x) for x in out[0]) == expected


def test_vectorised_matches_scalar():
    q = np.arange(10)
    vec = O.philox4x32(q, 3, 1, 0, 123, 456)
    for i in range(10):
        sc = O.philox4x32(i, 3, 1, 0, 123, 456)
        assert all(int(v[i]) == int(s) for v, s in zip(vec, sc, strict=True))


def test_c_matches_numpy_random_inputs():
    rng = np.random.default_rng(7)
    ctr = rng.integers(0, 2**32, size=(5000, 4), dtype=np.uint32)
    key = rng.integers(0, 2**32, size=(5000, 2), dtype=np.uint32)
    out = CO.philox(ctr, key)
    for i in range(0, 5000, 97):
        ref = O.philox4x32(*(int(x) for x in ctr[i]), int(key[i, 0]), int(key[i, 1]))
        assert [int(x) for x in ref] == [int(x) for x in out[i]]


def test_split_seed_and_site_stream_layout():
    assert O.split_seed(2**40 + 3) == (3, 256)
    assert O.split_seed(-1) == (0xFFFFFFFF, 0xFFFFFFFF)
    u = O.site_uniforms(16, 5, 1, 2**40 + 3)
    r = O.philox4x32(2, 5, 1, 0, 3, 256)
    assert [int(x) for x in u[8:12]] == [int(x) for x in r]


def test_site_uniforms_require_quads():
    with pytest.raises(ValueError):
        O.site_uniforms(6, 0, 0, 1)


def test_word_randoms_layout():
    lo, hi = O.word_randoms(8, 4, 1, 9, 2)
    r = O.philox4x32(5, 4, 1, 3, 9, 0)
    assert int(lo[5]) == int(r[0]) | (int(r[1]) << 32)
    assert int(hi[5]) == int(r[2]) | (int(r[3]) << 32)
    a0, a1 = O.word_randoms(8, 0, 0, 1, 0)
    b0, _ = O.word_randoms(8, 0, 0, 1, 1)
    assert a0.dtype == np.uint64
    assert not np.array_equal(a0, a1) and not np.array_equal(a0, b0)


def test_uniformity():
    u = O.site_uniforms(1 << 16, 0, 0, 1).astype(np.float64) / 2**32
    assert abs(u.mean() - 0.5) < 0.01
    hist, _ = np.histogram(u, bins=16, range=(0, 1))
    chi2 = ((hist - hist.mean()) ** 2 / hist.mean()).sum()
    assert chi2 < 50  # 15 degrees of freedom


def test_counter_truncates_to_32_bits():
    """Quad / word counters are taken mod 2^32 (the reference's (u32) casts)."""
    a = O.philox4x32(np.uint64(2**32 + 5), 1, 0, 0, 3, 4)
    b = O.philox4x32(5, 1, 0, 0, 3, 4)
    assert [int(x) for x in a] == [int(x) for x in b]


def test_planted_key_schedule_mistake_is_caught(monkeypatch):
    """A Philox that bumps the key before the first round fails the known answers."""
    real = O.philox4x32

    def wrong(c0, c1, c2, c3, k0, k1, rounds=10):
        return real(c0, c1, c2, c3, (k0 + O.W0) & O.MASK32, (k1 + O.W1) & O.MASK32, rounds)

    monkeypatch.setattr(O, "philox4x32", wrong)
    ctr, key, expected = KAT[2]
    assert tuple(int(x) for x in O.philox4x32(*ctr, *key)) != expected
