"""Pin the restated hot-start stream (oracle/pcg64.py) against numpy itself: raw outputs, the
jump-ahead, and the int8 draws of the reference's hot start; and (GPU) the device generator
against numpy for several seeds, sizes and offsets."""

import numpy as np
import pytest

from oracle import pcg64 as P

SEEDS = [0, 1, 7, 12345, 2**32 + 5, 2**63 - 1, 2**64 - 1]


@pytest.mark.parametrize("seed", SEEDS)
def test_raw_outputs_match_numpy(seed):
    assert P.raw(seed, 40) == [int(x) for x in np.random.PCG64(seed).random_raw(40)]


@pytest.mark.parametrize("delta", [0, 1, 2, 3, 1000, 2**20 + 17, 2**40 + 3])
def test_advance_matches_numpy(delta):
    bg = np.random.PCG64(99)
    s, inc = P.seed_state(99)
    bg.advance(delta)
    assert P.advance(s, inc, delta) == int(bg.state["state"]["state"])
    # and stepping one by one agrees with the jump for small deltas
    if delta <= 1000:
        t = s
        for _ in range(delta):
            t = P.step(t, inc)
        assert t == P.advance(s, inc, delta)


@pytest.mark.parametrize("seed", SEEDS)
@pytest.mark.parametrize("n", [8, 64, 1000, 4096])
def test_int8_draws_match_numpy(seed, n):
    ref = np.random.default_rng(seed).integers(0, 2, n, dtype=np.int8)
    np.testing.assert_array_equal(P.hot_start_bytes(seed, n), ref)


def test_planted_rotation_mistake_is_caught():
    """A rotate by (state >> 121) instead of >> 122 must fail the raw comparison."""
    s, inc = P.seed_state(3)
    s = P.step(s, inc)
    hi, lo = s >> 64, s & P.MASK64
    x, r = hi ^ lo, s >> 121 & 63
    wrong = ((x >> r) | (x << (64 - r))) & P.MASK64
    assert wrong != int(np.random.PCG64(3).random_raw(1)[0])


@pytest.mark.gpu
@pytest.mark.parametrize("seed", SEEDS)
@pytest.mark.parametrize("n", [8, 4096 * 8, 3 * 65536 + 24, 1 << 22])
def test_device_hot_start_matches_numpy(seed, n):
    from pyising_amd import _lib
    from pyising_amd.lattice import pcg64_state

    L = _lib.lib()
    buf = _lib.DeviceBuffer(n)
    st, inc = pcg64_state(seed)
    _lib.check(L.ising_hot_start(buf.ptr, n, st >> 64, st & P.MASK64, inc >> 64,
                                 inc & P.MASK64, None))
    got = np.empty(n, dtype=np.int8)
    buf.download(got)
    np.testing.assert_array_equal(got, np.random.default_rng(seed).integers(0, 2, n, dtype=np.int8))
