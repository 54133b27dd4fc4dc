"""Fused multi-spin-coded sweeps (many half sweeps per launch, ordered by per-tile completion
counters): bit-identical to the oracle and to one launch per half sweep, across launch chunks,
geometries of every supported row width, and the statistics pass. CPU: which geometries the
fused form accepts."""

import ctypes as C

import numpy as np
import pytest

from oracle import c_oracle as CO
from oracle import ising_oracle as O
from pyising_amd import _lib


def test_sync_workspace_size_needs_no_gpu():
    lib = _lib.load_library()
    n = C.c_size_t()
    assert lib.ising_msc_sync_bytes(16384, 128, C.byref(n)) == 0
    # 64 header words + one counter per 512-word tile, rounded to 256 bytes
    assert n.value == ((64 + 16384 * 128 // 512) * 4 + 255) // 256 * 256
    assert lib.ising_msc_sync_bytes(1, 128, C.byref(n)) == -1


def _sim(hip, shape, beta, seed, mode, lat):
    class Sim(hip.HipMultiSpin):
        pass

    Sim.mode = mode
    return Sim(shape, beta, seed=seed, init=lat)


@pytest.fixture(scope="module")
def hip():
    from pyising_amd.backends import hip_backend

    return hip_backend


@pytest.mark.gpu
@pytest.mark.parametrize("shape", [(512, 8192), (256, 16384), (128, 32768), (128, 65536),
                                   (1024, 16384)])
def test_fused_equals_oracle(hip, shape):
    """Every supported row width (nw = 64, 128, 256, 512 words), against the C oracle, over
    70 sweeps: three launches (32 + 32 + 6 sweeps) continue one counter workspace."""
    print(f"fused vs oracle {shape}", flush=True)
    lat = O.random_lattice(shape, 77)
    beta = 1 / 2.269
    sim = _sim(hip, shape, beta, 5, hip.MSC_FUSED, lat)
    assert sim.kernel()["mode"] == hip.MSC_FUSED
    ref = CO.CMsc(shape, beta, seed=5, init=lat)
    for n in (1, 69):
        sim.sweep(n)
        ref.sweep(n)
        assert sim.observables_exact() == tuple(ref.observables_exact())
        np.testing.assert_array_equal(sim.get_lattice(), ref.get_lattice())
    assert sim._epoch.value == 2 * 70
    sim.close()


@pytest.mark.gpu
@pytest.mark.parametrize("L", [8192, 16384])
def test_fused_equals_per_half_launches(hip, L):
    """The fused form and one launch per half sweep give the same words at the configured
    sizes, hot start at T_c, 100 sweeps (four launches of the fused form)."""
    print(f"fused vs per-half L={L}", flush=True)
    from pyising_amd.lattice import BETA_C_2D

    a = hip.HipMultiSpin((L, L), BETA_C_2D, seed=1)
    b = hip.HipMultiSpin((L, L), BETA_C_2D, seed=1)
    a.mode, b.mode = hip.MSC_FUSED, hip.MSC_COMPACT
    assert a.kernel()["mode"] == hip.MSC_FUSED and b.kernel()["mode"] == hip.MSC_COMPACT
    for n in (3, 97):
        a.sweep(n)
        b.sweep(n)
        assert a.observables_exact() == b.observables_exact()
        wa, wb = a.get_words(), b.get_words()
        assert all(np.array_equal(x, y) for x, y in zip(wa, wb))
    a.close()
    b.close()


@pytest.mark.gpu
def test_default_kernel_by_size(hip):
    """The default runs the fused form where a half sweep has at least 2048 tiles of 512 words
    and one launch per half sweep below (where the fused form would serialize)."""
    for shape, mode in (((16384, 16384), hip.MSC_FUSED), ((8192, 16384), hip.MSC_FUSED),
                        ((32768, 32768), hip.MSC_FUSED), ((8192, 8192), hip.MSC_COMPACT),
                        ((4096, 4096), hip.MSC_LOCKSTEP), ((11264, 11264), hip.MSC_COMPACT)):
        sim = hip.HipMultiSpin((128, 128), 0.4, seed=1, init="cold")
        sim.shape, sim.rows, sim.nw = shape, shape[0], shape[1] // 128
        assert sim.kernel()["mode"] == mode, shape


@pytest.mark.gpu
def test_fused_statistics_pass(hip):
    """The statistics variant of the fused kernel counts the same Philox calls as the
    per-half-sweep queue kernel (same tiles, same words)."""
    from pyising_amd.lattice import BETA_C_2D

    lat = O.random_lattice((2048, 8192), 3)
    a = _sim(hip, (2048, 8192), BETA_C_2D, 9, hip.MSC_FUSED, lat)
    b = _sim(hip, (2048, 8192), BETA_C_2D, 9, hip.MSC_COMPACT, lat)
    sa, sb = a.philox_stats(3), b.philox_stats(3)
    assert sa["words"] == sb["words"] == 3 * 2 * 2048 * 8192 // 128
    assert sa["calls_per_word"] == sb["calls_per_word"]
    np.testing.assert_array_equal(a.get_lattice(), b.get_lattice())


@pytest.mark.gpu
def test_fused_not_applicable_is_refused(hip):
    """Explicit fused mode on a geometry it does not cover is an error, not a silent
    fallback; the default picks another kernel there."""
    lat = O.random_lattice((64, 384), 1)
    sim = _sim(hip, (64, 384), 0.44, 1, hip.MSC_FUSED, lat)
    with pytest.raises(_lib.HipError, match="fused"):
        sim.sweep(1)
    d = hip.HipMultiSpin((64, 384), 0.44, seed=1, init=lat)
    assert d.kernel()["mode"] != hip.MSC_FUSED
