"""Resident sweeps (csrc/resident.hpp, int8_resident.hip, msc_resident.hip): a 2D lattice's half
sweeps in one cooperative launch, bands of rows kept in LDS with redundant halo rows, edge rows
exchanged between bands every few half sweeps. Every result is checked bit for bit against the
oracle's C restatement (pinned to the NumPy oracle and the reference's own vectors by the CPU
tests)."""

import numpy as np
import pytest

from oracle import c_oracle as CO
from oracle import ising_oracle as O


@pytest.fixture(scope="module")
def hip():
    from pyising_amd.backends import hip_backend

    return hip_backend


def _pair(hip, shape, beta, alg, seed, bands=0, mode=None):
    lat = O.random_lattice(shape, seed + 11)
    sim = hip.HipInt8(shape, beta, seed=seed, algorithm=alg, init=lat)
    sim.mode = hip.INT8_RESIDENT if mode is None else mode
    sim.resident_bands = bands
    ref = CO.CInt8(shape, beta, seed=seed, algorithm=alg, init=lat)
    return sim, ref


@pytest.mark.gpu
@pytest.mark.parametrize("shape,alg,beta", [
    ((64, 64), "metropolis", 0.4406868),
    ((128, 256), "heatbath", 1 / 2.3),
    ((256, 32), "metropolis", 0.5),        # lh = 16: one chunk per row
    ((1024, 1024), "metropolis", 0.4406868),
    ((512, 1024), "heatbath", 1 / 2.269),
    ((2048, 2048), "metropolis", 1 / 2.269),
    ((256, 4096), "heatbath", 1 / 2.269),  # 2048-byte rows: edge rows exchanged as rows
])
def test_resident_matches_oracle(hip, shape, alg, beta):
    """Calls of 1, 2, 3, 8 and 25 sweeps (rounds cut short, one round, many rounds), each
    compared with the oracle after the call, plus M and E."""
    sim, ref = _pair(hip, shape, beta, alg, seed=3)
    assert sim.resident_applies()
    for n in (1, 2, 3, 8, 25):
        sim.sweep(n)
        ref.sweep(n)
        np.testing.assert_array_equal(sim.get_lattice(), ref.get_lattice())
    assert sim.observables_exact() == tuple(ref.observables_exact())
    sim.close()


@pytest.mark.gpu
@pytest.mark.parametrize("bands", [8, 16, 32, 64, 128, 256])
def test_band_counts_change_nothing(hip, bands):
    """Every band count (rows per band from 128 down to 4, halo depth min(rows per band, 8))
    gives the oracle's lattice."""
    shape = (1024, 512)
    sim, ref = _pair(hip, shape, 0.44, "metropolis", seed=5, bands=bands)
    if not sim.resident_applies():
        sim.close()
        pytest.skip(f"{bands} bands do not fit this lattice on this device")
    sim.sweep(13)
    ref.sweep(13)
    np.testing.assert_array_equal(sim.get_lattice(), ref.get_lattice())
    sim.close()


@pytest.mark.gpu
def test_forms_interleave(hip):
    """Resident calls, one-launch-per-half-sweep calls and single half sweeps on the same
    lattice continue one stream; the default form picks resident sweeps for this size."""
    shape = (1024, 1024)
    sim, ref = _pair(hip, shape, 0.4406868, "metropolis", seed=9, mode=hip.INT8_DEFAULT)
    sim.sweep(4)
    assert sim._rs is not None  # the default used the resident form
    sim.mode = hip.INT8_HALVES
    sim.sweep(3)
    sim.mode = hip.INT8_RESIDENT
    sim.sweep(5)
    ref.sweep(12)
    np.testing.assert_array_equal(sim.get_lattice(), ref.get_lattice())
    sim.close()


@pytest.mark.gpu
def test_config_workload_L1024_100_sweeps(hip):
    """2D L = 1024, beta = 0.4406868, seed 1, hot start, 100 sweeps, then (M, E): the default
    (resident) form equals the one-launch-per-half-sweep form and the oracle's observables."""
    a = hip.HipInt8((1024, 1024), 0.4406868, seed=1)
    b = hip.HipInt8((1024, 1024), 0.4406868, seed=1)
    b.mode = hip.INT8_HALVES
    a.sweep(100)
    b.sweep(100)
    assert a._rs is not None and b._rs is None
    np.testing.assert_array_equal(a.get_lattice(), b.get_lattice())
    ref = CO.CInt8((1024, 1024), 0.4406868, seed=1, init=a.get_lattice())
    assert a.observables_exact() == b.observables_exact() == tuple(ref.observables_exact())
    a.close()
    b.close()


@pytest.mark.gpu
def test_not_applicable(hip):
    sim3 = hip.HipInt8((16, 16, 16), 0.22, seed=1)
    assert not sim3.resident_applies()
    sim3.mode = hip.INT8_RESIDENT
    with pytest.raises(ValueError):
        sim3.sweep(2)
    sim3.close()
    sim = hip.HipInt8((64, 24), 0.4, seed=1)  # lh = 12: not a multiple of 16
    assert not sim.resident_applies()
    sim.sweep(2)  # the default falls back to one launch per half sweep
    sim.close()


# ---------------------------------------------------------------------------------- msc ----
def _msc_pair(hip, shape, beta, seed, bands=0, mode=None):
    lat = O.random_lattice(shape, seed + 17)
    sim = hip.HipMultiSpin(shape, beta, seed=seed, init=lat)
    sim.mode = hip.MSC_RESIDENT if mode is None else mode
    sim.resident_bands = bands
    ref = CO.CMsc(shape, beta, seed=seed, init=lat)
    return sim, ref


@pytest.mark.gpu
@pytest.mark.parametrize("shape,beta", [
    ((64, 256), 0.44),          # nw = 2
    ((128, 512), 1 / 2.269),
    ((1024, 1024), 1 / 2.269),
    ((512, 2048), 0.5),         # nw = 16
    ((2048, 2048), 1 / 2.269),
    ((4096, 4096), 1 / 2.269),
])
def test_msc_resident_matches_oracle(hip, shape, beta):
    """Calls of 1, 2, 3, 8 and 25 sweeps, each compared with the oracle after the call, plus M
    and E."""
    sim, ref = _msc_pair(hip, shape, beta, seed=4)
    assert sim.resident_applies()
    for n in (1, 2, 3, 8, 25):
        sim.sweep(n)
        ref.sweep(n)
        np.testing.assert_array_equal(sim.get_lattice(), ref.get_lattice())
    assert sim.observables_exact() == tuple(ref.observables_exact())
    sim.close()


@pytest.mark.gpu
@pytest.mark.parametrize("bands", [8, 32, 128, 256])
def test_msc_band_counts_change_nothing(hip, bands):
    sim, ref = _msc_pair(hip, (1024, 1024), 0.44, seed=6, bands=bands)
    if not sim.resident_applies():
        sim.close()
        pytest.skip(f"{bands} bands do not fit this lattice on this device")
    sim.sweep(11)
    ref.sweep(11)
    np.testing.assert_array_equal(sim.get_lattice(), ref.get_lattice())
    sim.close()


@pytest.mark.gpu
def test_msc_forms_interleave(hip):
    """Resident calls and the other kernels on one lattice continue one stream; the default
    picks resident sweeps for this size; the eager ablation never does."""
    sim, ref = _msc_pair(hip, (2048, 2048), 1 / 2.269, seed=8, mode=hip.MSC_DEFAULT)
    sim.sweep(3)
    assert sim._rs is not None
    sim.mode = hip.MSC_COMPACT
    sim.sweep(2)
    sim.mode = hip.MSC_LOCKSTEP
    sim.sweep(1)
    sim.mode = hip.MSC_RESIDENT
    sim.sweep(4)
    ref.sweep(10)
    np.testing.assert_array_equal(sim.get_lattice(), ref.get_lattice())
    sim.close()
    eager = hip.HipMultiSpinEager((1024, 1024), 0.44, seed=1)
    eager.sweep(2)
    assert eager._rs is None
    eager.close()


@pytest.mark.gpu
def test_msc_not_applicable(hip):
    sim = hip.HipMultiSpin((64, 128), 0.44, seed=1)  # nw = 1: rows of 8 bytes
    assert not sim.resident_applies()
    sim.mode = hip.MSC_RESIDENT
    with pytest.raises(ValueError):
        sim.sweep(2)
    sim.mode = hip.MSC_DEFAULT
    sim.sweep(2)
    sim.close()
    big = hip.HipMultiSpin((16384, 16384), 0.44, seed=1)  # the fused sweeps' size
    big.sweep(2)
    assert big._rs is None
    big.close()


@pytest.mark.gpu
@pytest.mark.parametrize("variant", ["int8", "msc"])
def test_failure_bit_comes_back_with_observables(hip, variant):
    """A failure bit in the resident workspace, set from the host as a bounded wait that gave up
    sets it, is reported by the next observables_exact (read back with M and E), which zeroes
    the workspace; sweeps from a reset lattice then continue bit for bit. Every kernel ends
    normally: the waits see the bit and leave."""
    from pyising_amd import _lib

    if variant == "int8":
        sim, ref = _pair(hip, (1024, 1024), 0.44, "metropolis", seed=12)
    else:
        sim, ref = _msc_pair(hip, (1024, 1024), 0.44, seed=12)
    sim.sweep(2)
    sim.synchronize()
    _lib.check(sim._L.ising_fill_u64(sim._rs.ptr, 1, 1, sim.stream_ptr), "fill")
    sim.sweep(3)
    with pytest.raises(_lib.HipError, match=f"resident {variant} sweep failed"):
        sim.observables_exact()
    sim.set_lattice(ref.get_lattice())
    ref.t = sim.t
    sim.sweep(4)
    ref.sweep(4)
    np.testing.assert_array_equal(sim.get_lattice(), ref.get_lattice())
    assert sim.observables_exact() == tuple(ref.observables_exact())
    sim.close()
