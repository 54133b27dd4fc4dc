"""Pins the oracle's layouts, thresholds and half sweeps against independent restatements.

* layouts: round trips and the parity-grid selection (tests/test_lattice.py:19-30 of the
  reference);
* thresholds: the reference's own values at beta = 0.5 (tests/test_lattice.py:33-38), plus
  floor(p 2^32) recomputed with 60-digit decimal arithmetic at the workloads' betas;
* int8 sweeps: ``naive_sweep`` on the full lattice (tests/test_lattice.py:41-72);
* msc sweeps: the eager full-lattice statement (every 32-bit uniform assembled, int8 rule);
* observables: brute-force bond sums and the cold-start ground state;
* planted mistakes in the oracle (wrong horizontal neighbour, wrong level order) are caught.
"""

import math
from decimal import Decimal, getcontext

import numpy as np
import pytest

from oracle import ising_oracle as O
from pyising_amd import lattice as P


@pytest.mark.parametrize("shape", [(8, 16), (6, 8), (4, 6, 8), (2, 2, 8)])
def test_compressed_roundtrip(shape):
    lat = O.random_lattice(shape, 1)
    c0, c1 = O.to_compressed(lat)
    par = O.parity_grid(shape)
    np.testing.assert_array_equal(c0.ravel(), lat[par == 0])  # C order of colour-0 sites
    np.testing.assert_array_equal(c1.ravel(), lat[par == 1])
    np.testing.assert_array_equal(O.from_compressed(c0, c1, shape), lat)


def test_pack_roundtrip_and_bit_order():
    c = O.random_lattice((4, 256), 2)
    w = O.pack_bits(c)
    np.testing.assert_array_equal(O.unpack_bits(w), c)
    j = 64 * 1 + 5
    assert ((int(w[3, 1]) >> 5) & 1) == (1 if c[3, j] > 0 else 0)
    with pytest.raises(ValueError):
        O.pack_bits(np.ones((2, 96), np.int8))


def test_threshold_values_of_reference_tests():
    t = O.threshold_table(0.5, 2, "metropolis")
    assert t[0] == t[1] == t[2] == 0xFFFFFFFF
    assert t[3] == int(np.exp(-2.0) * 2**32) and t[4] == int(np.exp(-4.0) * 2**32)
    hb = O.threshold_table(0.5, 3, "heatbath")
    assert len(hb) == 7 and hb[3] == 2**31
    with pytest.raises(ValueError):
        O.threshold_table(0.5, 2, "glauber")


def _exact_threshold(p: Decimal) -> int:
    return min(int((p * (Decimal(2) ** 32)).to_integral_value(rounding="ROUND_FLOOR")), 2**32 - 1)


@pytest.mark.parametrize("beta,ndim", [(0.4406868, 2), (1 / P.TC_2D, 2), (1 / 1.5, 2),
                                       (1 / 2.269, 2), (1 / 3.0, 2), (0.221654626, 3),
                                       (0.20, 3), (0.24, 3), (0.42, 2), (0.22, 3)])
def test_thresholds_against_decimal_arithmetic(beta, ndim):
    getcontext().prec = 60
    b = Decimal(beta)
    z = 2 * ndim
    t = O.threshold_table(beta, ndim, "metropolis")
    h = O.threshold_table(beta, ndim, "heatbath")
    for i, x in enumerate(range(-z, z + 1, 2)):
        em = (-2 * b * x).exp()
        assert int(t[i]) == (0xFFFFFFFF if x <= 0 else _exact_threshold(em)), (beta, x)
        assert int(h[i]) == _exact_threshold(1 / (1 + em)), (beta, x)
    # the hot path's two msc thresholds at beta = 0.4406868
    if beta == 0.4406868:
        assert (int(t[3]), int(t[4])) == (736899868, 126432026)
    if beta == 1 / P.TC_2D:
        assert (int(t[3]), int(t[4])) == (736899888, 126432032)


@pytest.mark.parametrize("beta", [0.0, 0.3, 0.4406868, 0.7])
@pytest.mark.parametrize("ndim", [2, 3])
def test_product_tables_equal_oracle_tables(beta, ndim):
    for alg in ("metropolis", "heatbath"):
        np.testing.assert_array_equal(P.threshold_table(beta, ndim, alg),
                                      O.threshold_table(beta, ndim, alg))


def test_heatbath_symmetry():
    for beta in (0.1, 0.44, 0.9):
        h = O.threshold_table(beta, 2, "heatbath").astype(np.int64)
        assert np.all(np.abs(h + h[::-1] - 2**32) <= 1)
        assert np.all(np.diff(h) >= 0)


@pytest.mark.parametrize("shape", [(8, 16), (4, 4, 8), (16, 24), (6, 4, 8)])
@pytest.mark.parametrize("algorithm", ["metropolis", "heatbath"])
def test_int8_oracle_matches_naive(shape, algorithm):
    lat = O.random_lattice(shape, 3)
    sim = O.Int8Oracle(shape, 0.37, seed=42, algorithm=algorithm, init=lat)
    ref = lat
    for t in range(4):
        ref = O.naive_sweep(ref, 0.37, 42, t, algorithm)
    sim.sweep(4)
    np.testing.assert_array_equal(sim.get_lattice(), ref)
    assert sim.observables_exact() == O.observables_exact(ref)


@pytest.mark.parametrize("shape", [(16, 128), (64, 384), (6, 256)])
def test_msc_oracle_matches_eager_definition(shape):
    lat = O.random_lattice(shape, 9)
    sim = O.MscOracle(shape, 0.44, seed=11, init=lat)
    ref = lat
    for t in range(3):
        ref = O.msc_eager_sweep(ref, 0.44, 11, t)
    sim.sweep(3)
    np.testing.assert_array_equal(sim.get_lattice(), ref)
    assert sim.observables_exact() == O.observables_exact(ref)


def test_msc_at_beta_zero_matches_int8_rule():
    """At beta = 0 every threshold is UINT32_MAX: both streams flip every spin whose uniform is
    not 0xFFFFFFFF (tests/test_lattice.py:91-99 of the reference)."""
    lat = O.random_lattice((16, 128), 4)
    a = O.MscOracle((16, 128), 0.0, seed=5, init=lat)
    b = O.Int8Oracle((16, 128), 0.0, seed=5, init=lat)
    a.sweep(1)
    b.sweep(1)
    np.testing.assert_array_equal(a.get_lattice(), b.get_lattice())


def test_sweep_splitting_and_set_beta():
    a = O.Int8Oracle((16, 16), 0.4, seed=1)
    b = O.Int8Oracle((16, 16), 0.4, seed=1)
    a.sweep(5)
    b.sweep(2)
    b.sweep(3)
    np.testing.assert_array_equal(a.get_lattice(), b.get_lattice())
    m = O.MscOracle((16, 128), 0.3, seed=2)
    n = O.MscOracle((16, 128), 0.3, seed=2)
    m.sweep(2)
    m.set_beta(0.6)
    m.sweep(2)
    n.sweep(1)
    n.sweep(1)
    n.set_beta(0.6)
    n.sweep(2)
    np.testing.assert_array_equal(m.get_lattice(), n.get_lattice())


def test_cold_start_ground_state():
    assert O.MscOracle((16, 128), 0.5, init="cold").observables() == (1.0, -2.0)
    assert O.Int8Oracle((4, 4, 8), 0.5, init="cold").observables() == (1.0, -3.0)


def test_low_temperature_stays_ordered():
    sim = O.Int8Oracle((32, 32), 1.0, seed=1, init="cold")
    sim.sweep(50)
    assert abs(sim.observables()[0]) > 0.99


def test_shape_errors():
    for bad in [(8,), (8, 8, 8, 8), (7, 8), (8, 12)]:
        with pytest.raises(ValueError):
            O.check_shape(bad, 8)
    with pytest.raises(ValueError):
        O.MscOracle((16, 64), 0.3)


def test_hot_start_bits_equal_reference_hot_start():
    for shape, seed in [((16, 24), 1), ((4, 6, 8), 7), ((64, 128), 2**40 + 3)]:
        bits = P.hot_start_bits(shape, seed)
        np.testing.assert_array_equal(bits.astype(np.int8) * 2 - 1, O.random_lattice(shape, seed))


def test_planted_horizontal_neighbour_mistake_is_caught(monkeypatch):
    """An int8 oracle using j - 1 where the definition needs j + 1 disagrees with naive."""
    shape = (8, 16)
    lat = O.random_lattice(shape, 3)

    def wrong(self, color):
        o = self.c[1 - color].astype(np.int64)
        p1 = self.odd if color == 1 else ~self.odd  # parities swapped
        h = o + np.where(p1, np.roll(o, -1, axis=1), np.roll(o, 1, axis=1))
        g = o.reshape(self.shape[:-1] + (self.lh,))
        h += (np.roll(g, 1, axis=0) + np.roll(g, -1, axis=0)).reshape(self.rows, self.lh)
        return h

    monkeypatch.setattr(O.Int8Oracle, "neighbor_sum", wrong)
    sim = O.Int8Oracle(shape, 0.37, seed=42, init=lat)
    sim.sweep(2)
    ref = O.naive_sweep(O.naive_sweep(lat, 0.37, 42, 0, "metropolis"), 0.37, 42, 1, "metropolis")
    assert not np.array_equal(sim.get_lattice(), ref)


def test_planted_level_order_mistake_is_caught(monkeypatch):
    """An msc oracle that swaps the two 64-bit levels of a Philox call disagrees with eager."""
    shape = (16, 128)
    lat = O.random_lattice(shape, 9)
    real = O.word_randoms
    monkeypatch.setattr(O, "word_randoms", lambda *a: real(*a)[::-1])
    sim = O.MscOracle(shape, 0.44, seed=11, init=lat)
    sim.sweep(1)
    monkeypatch.setattr(O, "word_randoms", real)
    ref = O.msc_eager_sweep(lat, 0.44, 11, 0)
    assert not np.array_equal(sim.get_lattice(), ref)


def test_math_exp_is_what_the_tables_use():
    assert O.prob_to_threshold(math.exp(-2.0)) == int(np.exp(-2.0) * 2**32)
    assert O.prob_to_threshold(1.0) == O.prob_to_threshold(2.0) == 0xFFFFFFFF
