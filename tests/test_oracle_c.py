"""The C restatement of the oracle equals the NumPy restatement bit for bit (CPU only), so it
can stand in for it on large and long runs (the GPU parity tests at full sizes)."""

import numpy as np
import pytest

from oracle import c_oracle as CO
from oracle import ising_oracle as O


@pytest.mark.parametrize("shape", [(16, 24), (8, 8, 16), (64, 32), (4, 6, 8), (2, 8), (2, 2, 8)])
@pytest.mark.parametrize("algorithm", ["metropolis", "heatbath"])
def test_int8_c_equals_numpy(shape, algorithm):
    lat = O.random_lattice(shape, 5)
    beta = 0.42 if len(shape) == 2 else 0.22
    a = O.Int8Oracle(shape, beta, seed=2**40 + 7, algorithm=algorithm, init=lat)
    b = CO.CInt8(shape, beta, seed=2**40 + 7, algorithm=algorithm, init=lat)
    for _ in range(3):
        a.sweep(2)
        b.sweep(2)
        assert a.observables_exact() == b.observables_exact()
    np.testing.assert_array_equal(a.get_lattice(), b.get_lattice())


@pytest.mark.parametrize("shape", [(128, 128), (64, 384), (16, 1024), (130, 256), (2, 128)])
@pytest.mark.parametrize("beta", [0.0, 0.44, 1.2])
def test_msc_c_equals_numpy(shape, beta):
    lat = O.random_lattice(shape, 9)
    a = O.MscOracle(shape, beta, seed=11, init=lat)
    b = CO.CMsc(shape, beta, seed=11, lattice=lat)
    a.sweep(4)
    b.sweep(4)
    np.testing.assert_array_equal(a.w[0], b.w[0])
    np.testing.assert_array_equal(a.w[1], b.w[1])
    assert a.observables_exact() == b.observables_exact()


def test_thread_count_does_not_change_results():
    shape = (64, 256)
    runs = []
    for n in (1, 3, 8):
        CO.set_threads(n)
        s = CO.CMsc(shape, 0.44, seed=3)
        s.sweep(5)
        i = CO.CInt8(shape, 0.44, seed=3)
        i.sweep(5)
        runs.append((s.w[0].copy(), i.get_lattice()))
    CO.set_threads(8)
    for w, lat in runs[1:]:
        np.testing.assert_array_equal(w, runs[0][0])
        np.testing.assert_array_equal(lat, runs[0][1])


def test_msc_call_count_at_tc_matches_published_mean():
    """Mean Philox calls per word at T_c after equilibration: the reference measured 3.804
    (results/rtx3070-laptop.json:21 of the reference)."""
    s = CO.CMsc((512, 512), 1 / (2 / np.log(1 + np.sqrt(2))), seed=1)
    s.sweep(200)
    s.calls[:] = 0
    s.sweep(4)
    cpw = int(s.calls[0]) / (2 * 4 * s.rows * s.nw)
    assert 3.6 < cpw < 4.0
