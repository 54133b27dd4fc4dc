"""Host statistics (pyising_amd.observables): closed forms, an independent leave-one-block-out
restatement of the jackknife, crossing interpolation, sampling, and (GPU) the batched scan's
block-moment jackknife against per-sweep time series of the same lattices."""

import numpy as np
import pytest

from pyising_amd import observables as obs


def test_binder_limits():
    assert obs.binder(np.array([1.0, -1.0, 1.0, 1.0])) == pytest.approx(2 / 3)
    rng = np.random.default_rng(0)
    assert abs(obs.binder(rng.standard_normal(400_000))) < 0.01  # Gaussian: U_4 = 0


def test_binning_error_iid():
    rng = np.random.default_rng(1)
    x = rng.standard_normal(64_000)
    err = obs.binning_error(x)
    assert 0.7 < err * np.sqrt(x.size) < 1.3
    assert np.isnan(obs.binning_error(np.array([1.0])))
    # a constant series has no spread
    assert obs.binning_error(np.full(100, 3.0)) == 0.0


def test_jackknife_matches_leave_one_block_out():
    rng = np.random.default_rng(2)
    m = np.tanh(rng.standard_normal(3203))  # not a multiple of 32: the tail is dropped
    nb = 32
    per = m.size // nb
    u_i = []
    for i in range(nb):
        rest = np.delete(m[: nb * per].reshape(nb, per), i, axis=0).ravel()
        u_i.append(1 - np.mean(rest**4) / (3 * np.mean(rest**2) ** 2))
    u_i = np.array(u_i)
    expect = np.sqrt((nb - 1) / nb * np.sum((u_i - u_i.mean()) ** 2))
    assert obs.jackknife_binder_error(m) == pytest.approx(expect, rel=1e-12)
    u, err = obs.binder_from_blocks((m[: nb * per].reshape(nb, per) ** 2).mean(1),
                                    (m[: nb * per].reshape(nb, per) ** 4).mean(1))
    assert err == pytest.approx(expect, rel=1e-12)
    assert u == pytest.approx(obs.binder(m[: nb * per]), rel=1e-12)


def test_crossing():
    T = np.linspace(2.0, 2.5, 6)
    a = 1.0 - (T - 2.0)  # crosses b at T = 2.2692...
    b = 0.5 + 0.8 * (T - 2.0) - 0.2
    t_x = obs.crossing(T, a, b)
    assert t_x == pytest.approx(2.0 + 0.7 / 1.8, rel=1e-12)
    assert obs.crossing(T, a, a + 1) != obs.crossing(T, a, a + 1)  # NaN: no crossing
    assert obs.crossing(T, b, b) == T[0]  # identical curves: first grid point


class _Fake:
    """A deterministic 'simulation' whose observables are its sweep count."""

    def __init__(self):
        self.t = 0

    def sweep(self, n):
        self.t += n

    def observables(self):
        return (self.t / 100.0, -self.t / 50.0)


def test_sample_and_summary():
    s = obs.sample(_Fake(), n_therm=10, n_meas=5, every=3)
    np.testing.assert_allclose(s.m, [0.13, 0.16, 0.19, 0.22, 0.25])
    np.testing.assert_allclose(s.e, -2 * s.m)
    summ = s.summary()
    assert set(summ) == {"abs_m", "abs_m_err", "e", "e_err", "m2", "m4", "binder", "binder_err", "n"}
    assert summ["n"] == 5 and summ["abs_m"] == pytest.approx(0.19)


class _FakeSeries(_Fake):
    """The same 'simulation' offering a device-side series of exact (M, E) over 100 spins."""

    n_spins = 100

    def measure_series(self, n, every):
        out = []
        for _ in range(n):
            self.sweep(every)
            out.append((self.t, -2 * self.t))
        return np.array(out, dtype=np.int64)


def test_sample_uses_the_device_series():
    a = obs.sample(_Fake(), n_therm=10, n_meas=5, every=3)
    b = obs.sample(_FakeSeries(), n_therm=10, n_meas=5, every=3)
    np.testing.assert_allclose(b.m, a.m)
    np.testing.assert_allclose(b.e, a.e)


@pytest.mark.gpu
@pytest.mark.parametrize("variant,shape", [("int8", (64, 64)), ("int8", (16, 16, 16)),
                                           ("msc", (128, 256))])
def test_measure_series_equals_the_host_loop(variant, shape):
    from pyising_amd.backends.hip_backend import HipInt8, HipMultiSpin

    cls = HipInt8 if variant == "int8" else HipMultiSpin
    beta = 0.22 if len(shape) == 3 else 0.44
    a, b = cls(shape, beta, seed=5), cls(shape, beta, seed=5)
    ser = a.measure_series(7, every=3)
    loop = []
    for _ in range(7):
        b.sweep(3)
        loop.append(b.observables_exact())
    np.testing.assert_array_equal(ser, np.array(loop, dtype=np.int64))
    assert a.t == b.t == 21


@pytest.mark.gpu
def test_scan_binder_equals_time_series():
    from pyising_amd.backends.hip_backend import HipInt8, HipMultiSpin
    from pyising_amd.batch import BatchScan

    specs = [(HipInt8, (32, 32), 1 / 2.3, 11), (HipMultiSpin, (128, 256), 1 / 2.25, 12),
             (HipInt8, (8, 8, 16), 0.23, 13)]
    sims = [cls(shape, beta, seed=seed) for cls, shape, beta, seed in specs]
    refs = [cls(shape, beta, seed=seed) for cls, shape, beta, seed in specs]
    scan = BatchScan(sims)
    n_blocks, per = 6, 4
    got = obs.scan_binder(scan, n_blocks, per)
    scan.close()
    for ref, (u, err) in zip(refs, got, strict=True):
        m = []
        for _ in range(n_blocks * per):
            ref.sweep(1)
            m.append(ref.observables()[0])
        m = np.array(m)
        u_ref, err_ref = obs.binder_from_blocks((m.reshape(n_blocks, per) ** 2).mean(1),
                                                (m.reshape(n_blocks, per) ** 4).mean(1))
        assert u == pytest.approx(u_ref, rel=1e-12, abs=1e-12)
        assert err == pytest.approx(err_ref, rel=1e-9, abs=1e-12)
