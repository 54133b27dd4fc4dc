"""The hip backend reproduces the oracle bit for bit (same Philox stream, same thresholds).

Mirrors the reference's GPU identity suite (tests/test_backends.py:9-95 of the reference):
int8 identity over shapes x algorithms, msc identity, set_beta with split sweeps, hot and cold
starts; adds config-shaped cases, exact integer observables and every msc kernel mode.
"""

import numpy as np
import pytest

from oracle import c_oracle as CO
from oracle import ising_oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def hip():
    from pyising_amd.backends import hip_backend

    return hip_backend


@pytest.mark.parametrize("shape", [(16, 24), (8, 8, 16), (64, 32), (32, 64), (4, 6, 8)])
@pytest.mark.parametrize("algorithm", ["metropolis", "heatbath"])
def test_int8_bitwise_identical(hip, shape, algorithm):
    lat = O.random_lattice(shape, 5)
    beta = 0.42 if len(shape) == 2 else 0.22
    ref = O.Int8Oracle(shape, beta, seed=2**40 + 7, algorithm=algorithm, init=lat)
    sim = hip.HipInt8(shape, beta, seed=2**40 + 7, algorithm=algorithm, init=lat)
    ref.sweep(6)
    sim.sweep(6)
    np.testing.assert_array_equal(sim.get_lattice(), ref.get_lattice())
    assert sim.observables_exact() == ref.observables_exact()
    m, e = sim.observables()
    assert m == pytest.approx(O.observables_exact(ref.get_lattice())[0] / lat.size)
    assert e == pytest.approx(O.observables_exact(ref.get_lattice())[1] / lat.size)


# int8 per-half-sweep kernel layouts (ising_int8_kernel): segment waves (a wave is whole rows of
# one plane), row waves (a wave's lanes in one row), per-lane rows; with the sites per lane the
# host picks for each shape, and both parities of the first row (odd row counts included)
INT8_LAYOUT_SHAPES = [((64, 64), 2), ((130, 256), 2), ((96, 128), 2), ((256, 4096), 1),
                      ((512, 2048), 1), ((32, 8, 64), 2), ((6, 96, 256), 2), ((16, 16, 16), 0),
                      ((10, 12, 16), 0)]


@pytest.mark.parametrize("shape,layout", INT8_LAYOUT_SHAPES)
@pytest.mark.parametrize("algorithm", ["metropolis", "heatbath"])
def test_int8_half_sweep_layouts(hip, shape, layout, algorithm):
    """One launch per half sweep (not the resident sweeps) in every wave layout, against the C
    oracle after 1 and 7 sweeps."""
    print(f"int8 layout {layout} {shape} {algorithm}", flush=True)
    lat = O.random_lattice(shape, 21)
    beta = 0.43 if len(shape) == 2 else 0.222
    sim = hip.HipInt8(shape, beta, seed=2**33 + 5, algorithm=algorithm, init=lat)
    sim.mode = hip.INT8_HALVES
    assert sim.kernel()["layout"] == layout
    ref = CO.CInt8(shape, beta, seed=2**33 + 5, algorithm=algorithm, init=lat)
    for n in (1, 6):
        sim.sweep(n)
        ref.sweep(n)
        np.testing.assert_array_equal(sim.get_lattice(), ref.get_lattice())
        assert sim.observables_exact() == tuple(ref.observables_exact())
    sim.close()


MSC_MODES = [(0, 42), (0, 32), (0, 52), (0, 62), (0, 44), (0, 34), (1, 0), (2, 0)]


@pytest.mark.parametrize("mode,k", MSC_MODES)
# classification layouts: rows of nw = Lx/128 words; nw % 64 == 0 -> whole-row waves
# ((16, 8192), (8, 16384)); nw dividing 64 with a multiple of 64 words -> row-segment waves
# ((128, 128), (16, 1024), (256, 256), (64, 512), (64, 2048), (32, 4096)); otherwise per-lane
# loads ((64, 384), (130, 256), (66, 128), (6, 24576))
@pytest.mark.parametrize("shape", [(128, 128), (64, 384), (16, 1024), (130, 256),
                                   (16, 8192), (8, 16384), (6, 24576), (256, 256), (64, 512),
                                   (64, 2048), (32, 4096), (66, 128)])
def test_msc_bitwise_identical(hip, shape, mode, k):
    lat = O.random_lattice(shape, 9)
    ref = O.MscOracle(shape, 0.44, seed=11, init=lat)
    cls = type("M", (hip.HipMultiSpin,), {"mode": mode, "tile_k": k or 42})
    sim = cls(shape, 0.44, seed=11, init=lat)
    ref.sweep(5)
    sim.sweep(5)
    np.testing.assert_array_equal(sim.get_lattice(), ref.get_lattice())
    assert sim.observables_exact() == ref.observables_exact()


@pytest.mark.parametrize("variant", ["int8", "msc"])
def test_set_beta_and_split_sweeps(hip, variant):
    cls = hip.HipInt8 if variant == "int8" else hip.HipMultiSpin
    shape = (128, 128)
    lat = O.random_lattice(shape, 1)
    a = cls(shape, 0.3, seed=3, init=lat)
    b = cls(shape, 0.3, seed=3, init=lat)
    a.sweep(2)
    a.set_beta(0.5)
    a.sweep(3)
    b.sweep(2)
    b.set_beta(0.5)
    b.sweep(1)
    b._half_sweep(0)
    b._half_sweep(1)
    b.t += 1
    b.sweep(1)
    np.testing.assert_array_equal(a.get_lattice(), b.get_lattice())
    ref = (O.Int8Oracle if variant == "int8" else O.MscOracle)(shape, 0.3, seed=3, init=lat)
    ref.sweep(2)
    ref.set_beta(0.5)
    ref.sweep(3)
    np.testing.assert_array_equal(a.get_lattice(), ref.get_lattice())


@pytest.mark.parametrize("variant", ["int8", "msc", "msc-eager32"])
def test_hot_and_cold_starts(variant):
    from pyising_amd import create

    shape = (128, 128)
    cold = create("hip", variant, shape=shape, beta=0.5, init="cold")
    assert cold.observables() == pytest.approx((1.0, -2.0))
    assert cold.observables_exact() == (shape[0] * shape[1], -2 * shape[0] * shape[1])
    hot = create("hip", variant, shape=shape, beta=0.5, seed=1, init="hot")
    np.testing.assert_array_equal(hot.get_lattice(), O.random_lattice(shape, 1))
    m, e = hot.observables()
    assert abs(m) < 0.05 and abs(e) < 0.05


@pytest.mark.parametrize("variant,shape", [("int8", (64, 48)), ("int8", (8, 6, 16)),
                                           ("msc", (64, 256)), ("msc", (130, 384))])
@pytest.mark.parametrize("seed", [0, 7, 2**63 + 5, 2**64 - 1])
def test_device_hot_start_equals_numpy_draws(variant, shape, seed):
    """The device-generated hot start is numpy default_rng(seed)'s draws mapped 0 -> -1."""
    from pyising_amd import create

    sim = create("hip", variant, shape=shape, beta=0.4, seed=seed, init="hot")
    np.testing.assert_array_equal(sim.get_lattice(), O.random_lattice(shape, seed))


@pytest.mark.parametrize("variant,shape", [("int8", (4096, 8192)), ("msc", (1024, 8192))])
def test_device_hot_start_large(variant, shape):
    """Hot starts larger than one generation of layout workgroups."""
    from pyising_amd import create

    sim = create("hip", variant, shape=shape, beta=0.4, seed=3, init="hot")
    np.testing.assert_array_equal(sim.get_lattice(), O.random_lattice(shape, 3))


def test_msc_131072_set_up_covers_every_word(hip):
    """A 131072^2 msc lattice (2^27 words per colour: more work-items than one dispatch may
    hold) is set up whole: the cold start gives M = N, E = -2N exactly, and the hot start's M
    and E are within 6 sigma of 0 (independent fair spins: var M = N, var E = 2N)."""
    shape = (131072, 131072)
    n = shape[0] * shape[1]
    cold = hip.HipMultiSpin(shape, 1 / 2.269, seed=1, init="cold")
    assert cold.observables_exact() == (n, -2 * n)
    del cold
    hot = hip.HipMultiSpin(shape, 1 / 2.269, seed=1, init="hot")
    m, e = hot.observables_exact()
    assert abs(m) < 6 * n**0.5 and abs(e) < 6 * (2 * n) ** 0.5, (m, e)
    hot.sweep(1)
    m1, e1 = hot.observables_exact()
    assert (m1, e1) != (m, e) and e1 < -0.5 * n  # one sweep at T_c orders most bonds


@pytest.mark.parametrize("variant,shape", [("int8", (64, 96)), ("int8", (8, 8, 16)),
                                           ("msc", (128, 256))])
def test_resume_from_checkpoint(hip, variant, shape):
    """A checkpoint is the lattice and the sweep counter: a new simulation set to them continues
    bit-identically (the counter-based stream needs nothing else)."""
    cls = hip.HipInt8 if variant == "int8" else hip.HipMultiSpin
    beta = 0.2216 if len(shape) == 3 else 0.44
    a = cls(shape, beta, seed=77)
    a.sweep(5)
    lat, t = a.get_lattice(), a.t
    b = cls(shape, beta, seed=77, init=lat)
    b.t = t
    a.sweep(3)
    b.sweep(3)
    np.testing.assert_array_equal(a.get_lattice(), b.get_lattice())
    assert a.observables_exact() == b.observables_exact()


def test_cold_start_3d(hip):
    sim = hip.HipInt8((8, 8, 16), 0.3, init="cold")
    assert sim.observables_exact() == (1024, -3 * 1024)


# the last three exceed the layout kernels' 8192-block grid, so their grid-stride loops wrap
@pytest.mark.parametrize("variant,shape", [("int8", (16, 24)), ("int8", (6, 4, 8)),
                                           ("msc", (64, 384)), ("int8", (4096, 8192)),
                                           ("int8", (64, 64, 8192)), ("msc", (1024, 8192))])
def test_layout_roundtrip(hip, variant, shape):
    cls = hip.HipInt8 if variant == "int8" else hip.HipMultiSpin
    lat = O.random_lattice(shape, 123)
    sim = cls(shape, 0.4, seed=1, init=lat)
    np.testing.assert_array_equal(sim.get_lattice(), lat)
    if variant == "msc":
        c0, c1 = O.to_compressed(lat)
        w0, w1 = sim.get_words()
        np.testing.assert_array_equal(w0, O.pack_bits(c0))
        np.testing.assert_array_equal(w1, O.pack_bits(c1))


def test_philox_known_answers():
    import ctypes as C

    from pyising_amd import _lib

    ctr = np.array([[0, 0, 0, 0], [0xFFFFFFFF] * 4,
                    [0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344]], dtype=np.uint32)
    key = np.array([[0, 0], [0xFFFFFFFF] * 2, [0xA4093822, 0x299F31D0]], dtype=np.uint32)
    out = np.zeros_like(ctr)
    _lib.check(_lib.lib().ising_philox_kat(ctr.ctypes.data, key.ctypes.data, out.ctypes.data, 3))
    exp = np.array([[0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8],
                    [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD],
                    [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]], dtype=np.uint32)
    np.testing.assert_array_equal(out, exp)
    rng = np.random.default_rng(0)
    ctr = rng.integers(0, 2**32, size=(4096, 4), dtype=np.uint32)
    key = rng.integers(0, 2**32, size=(4096, 2), dtype=np.uint32)
    out = np.zeros_like(ctr)
    _lib.check(_lib.lib().ising_philox_kat(ctr.ctypes.data, key.ctypes.data, out.ctypes.data,
                                           4096))
    np.testing.assert_array_equal(out, CO.philox(ctr, key))
    del C


# ----------------------------------------------------------- config-shaped cases (C oracle) --
def test_config_2d_int8_L1024_100_sweeps(hip):
    """2D int8 L=1024, beta=0.4406868, seed=1, hot start, 100 sweeps, then (M, E)."""
    shape, beta = (1024, 1024), 0.4406868
    sim = hip.HipInt8(shape, beta, seed=1)
    ref = CO.CInt8(shape, beta, seed=1)
    sim.sweep(100)
    ref.sweep(100)
    np.testing.assert_array_equal(sim.get_lattice(), ref.get_lattice())
    assert sim.observables_exact() == ref.observables_exact()


@pytest.mark.parametrize("algorithm", ["metropolis", "heatbath"])
@pytest.mark.parametrize("K,init", [(0.221654626, "hot"), (0.20, "cold"), (0.24, "hot")])
def test_config_3d_int8(hip, algorithm, K, init):
    """3D int8, seed=7, cold and hot, (E, |M|) every 10 sweeps (L=32 here)."""
    shape = (32, 32, 32)
    sim = hip.HipInt8(shape, K, seed=7, algorithm=algorithm, init=init)
    ref = CO.CInt8(shape, K, seed=7, algorithm=algorithm, init=init)
    for _ in range(4):
        sim.sweep(10)
        ref.sweep(10)
        assert sim.observables_exact() == ref.observables_exact()
    np.testing.assert_array_equal(sim.get_lattice(), ref.get_lattice())


def test_config_msc_L1024_tc(hip):
    from pyising_amd import BETA_C_2D

    shape = (1024, 1024)
    sim = hip.HipMultiSpin(shape, BETA_C_2D, seed=1)
    ref = CO.CMsc(shape, BETA_C_2D, seed=1)
    sim.sweep(50)
    ref.sweep(50)
    w = sim.get_words()
    np.testing.assert_array_equal(w[0], ref.w[0])
    np.testing.assert_array_equal(w[1], ref.w[1])
    assert sim.observables_exact() == ref.observables_exact()


@pytest.mark.parametrize("T", [1.5, 2.269, 3.0])
def test_config_int8_L2048_temperatures(hip, T):
    shape = (2048, 2048)
    sim = hip.HipInt8(shape, 1 / T, seed=1)
    ref = CO.CInt8(shape, 1 / T, seed=1)
    sim.sweep(5)
    ref.sweep(5)
    assert sim.observables_exact() == ref.observables_exact()
    np.testing.assert_array_equal(sim.get_lattice(), ref.get_lattice())


@pytest.mark.parametrize("T", [1.5, 2.269, 3.0])
def test_config_int8_L8192_temperatures(hip, T):
    """2D int8 Metropolis at the benchmark size, L=8192, T in {1.5, 2.269, 3.0}, seed=1, hot."""
    shape = (8192, 8192)
    sim = hip.HipInt8(shape, 1 / T, seed=1)
    ref = CO.CInt8(shape, 1 / T, seed=1)
    sim.sweep(4)
    ref.sweep(4)
    assert sim.observables_exact() == ref.observables_exact()
    np.testing.assert_array_equal(sim.get_lattice(), ref.get_lattice())


@pytest.mark.parametrize("algorithm", ["metropolis", "heatbath"])
@pytest.mark.parametrize("K,init", [(0.221654626, "hot"), (0.221654626, "cold"), (0.20, "cold"),
                                    (0.24, "hot")])
def test_config_3d_int8_L256(hip, algorithm, K, init):
    """3D int8 at the configuration's own size, L=256 (16.8 M spins), seed=7: (M, E) every 10
    sweeps and the lattice after 30."""
    shape = (256, 256, 256)
    sim = hip.HipInt8(shape, K, seed=7, algorithm=algorithm, init=init)
    ref = CO.CInt8(shape, K, seed=7, algorithm=algorithm, init=init)
    for _ in range(3):
        sim.sweep(10)
        ref.sweep(10)
        assert sim.observables_exact() == ref.observables_exact()
    np.testing.assert_array_equal(sim.get_lattice(), ref.get_lattice())


# sites per lane follow the lattice size (the widest of 16/8/4 dividing Lx/2 that gives >= 2^17
# lanes): (2048, 1024) and (64, 128, 256) take 8, (128, 128, 256) takes 16, (32, 32, 32) and
# (1024, 1024) take 4, (2048, 2048) takes 16; every width must give the same spins
@pytest.mark.parametrize("shape", [(2048, 1024), (64, 128, 256), (128, 128, 256)])
@pytest.mark.parametrize("algorithm", ["metropolis", "heatbath"])
def test_int8_every_vector_width(hip, shape, algorithm):
    beta = 0.44 if len(shape) == 2 else 0.2216
    sim = hip.HipInt8(shape, beta, seed=99, algorithm=algorithm)
    ref = CO.CInt8(shape, beta, seed=99, algorithm=algorithm)
    sim.sweep(3)
    ref.sweep(3)
    assert sim.observables_exact() == ref.observables_exact()
    np.testing.assert_array_equal(sim.get_lattice(), ref.get_lattice())


def test_config_msc_L16384_tc_few_sweeps(hip):
    """The headline lattice itself: msc L=16384 at T_c, seed=1, hot start, 3 sweeps."""
    from pyising_amd import BETA_C_2D

    shape = (16384, 16384)
    sim = hip.HipMultiSpin(shape, BETA_C_2D, seed=1)
    ref = CO.CMsc(shape, BETA_C_2D, seed=1)
    sim.sweep(3)
    ref.sweep(3)
    w0, w1 = sim.get_words()
    assert np.array_equal(w0, ref.w[0]) and np.array_equal(w1, ref.w[1])
    assert sim.observables_exact() == ref.observables_exact()


def _random_cases(n=36, seed=2024):
    """Seeded draws over the shapes the layouts allow: int8 2D/3D (every dimension even, lx a
    multiple of 8) with both algorithms, msc 2D (lx a multiple of 128) in every kernel mode."""
    rng = np.random.default_rng(seed)
    cases = []
    for i in range(n):
        kind = ("int8_2d", "int8_3d", "msc")[i % 3]
        beta = float(rng.uniform(0.05, 0.8))
        seed_i = int(rng.integers(0, 2**63))
        if kind == "int8_2d":
            shape = (2 * int(rng.integers(1, 40)), 8 * int(rng.integers(1, 40)))
            alg = ("metropolis", "heatbath")[int(rng.integers(0, 2))]
            cases.append((kind, shape, beta, seed_i, alg, None))
        elif kind == "int8_3d":
            shape = (2 * int(rng.integers(1, 6)), 2 * int(rng.integers(1, 8)),
                     8 * int(rng.integers(1, 5)))
            alg = ("metropolis", "heatbath")[int(rng.integers(0, 2))]
            cases.append((kind, shape, beta * 0.5, seed_i, alg, None))
        else:
            shape = (2 * int(rng.integers(1, 80)), 128 * int(rng.integers(1, 70)))
            mode = [(0, 42), (0, 32), (1, 0), (2, 0), (3, 42)][int(rng.integers(0, 5))]
            cases.append((kind, shape, beta, seed_i, "metropolis", mode))
    return cases


@pytest.mark.parametrize("case", _random_cases(), ids=lambda c: f"{c[0]}-{'x'.join(map(str, c[1]))}")
def test_random_shapes_match_oracle(hip, case):
    kind, shape, beta, seed, alg, mode = case
    lat = O.random_lattice(shape, seed % 1000)
    if kind == "msc":
        ref = O.MscOracle(shape, beta, seed=seed, init=lat)
        cls = type("M", (hip.HipMultiSpin,), {"mode": mode[0], "tile_k": mode[1] or 42})
        sim = cls(shape, beta, seed=seed, init=lat)
    else:
        ref = O.Int8Oracle(shape, beta, seed=seed, algorithm=alg, init=lat)
        sim = hip.HipInt8(shape, beta, seed=seed, algorithm=alg, init=lat)
    ref.sweep(3)
    sim.sweep(3)
    np.testing.assert_array_equal(sim.get_lattice(), ref.get_lattice())
    assert sim.observables_exact() == ref.observables_exact()
