"""Batched temperature scans: every lattice of a batched launch (one launch per half sweep, or
dependency-ordered launches of many sweeps) is bit-identical to the same lattice swept alone and
to the oracle, and the device-side Binder sums are exact."""

import ctypes as C

import numpy as np
import pytest

from oracle import c_oracle as CO
from oracle import ising_oracle as O


def _mixed(hip, seed0=100):
    """A batch covering every tile kind: int8 2D/3D x Metropolis/heat bath, msc of several
    word counts, uneven sizes and temperatures."""
    specs = [
        (hip.HipInt8, (64, 64), 1 / 2.2, "metropolis"),
        (hip.HipInt8, (32, 48), 1 / 2.3, "heatbath"),
        (hip.HipInt8, (16, 24), 0.42, "metropolis"),  # V = 4
        (hip.HipInt8, (8, 8, 16), 0.22, "metropolis"),
        (hip.HipInt8, (6, 4, 8), 0.24, "heatbath"),
        (hip.HipMultiSpin, (128, 256), 0.44, "metropolis"),
        (hip.HipMultiSpin, (64, 384), 1 / 2.25, "metropolis"),
        (hip.HipMultiSpin, (130, 128), 0.5, "metropolis"),
        (hip.HipInt8, (256, 256), 1 / 2.269, "metropolis"),
        (hip.HipMultiSpin, (1024, 1024), 1 / 2.269, "metropolis"),
    ]
    out = []
    for i, (cls, shape, beta, alg) in enumerate(specs):
        lat = O.random_lattice(shape, 7 + i)
        out.append((cls, shape, beta, alg, seed0 + i, lat))
    return out


def _make(spec):
    cls, shape, beta, alg, seed, lat = spec
    return cls(shape, beta, seed=seed, algorithm=alg, init=lat)


def _oracle(spec):
    """The same lattice in the oracle's C restatement (pinned to the NumPy oracle and to the
    reference's own vectors by the CPU tests)."""
    cls, shape, beta, alg, seed, lat = spec
    ocls = CO.CMsc if cls.variant == "msc" else CO.CInt8
    return ocls(shape, beta, seed=seed, algorithm=alg, init=lat)


@pytest.fixture(scope="module")
def hip():
    from pyising_amd.backends import hip_backend

    return hip_backend


@pytest.mark.gpu
@pytest.mark.parametrize("persistent", [False, True])
@pytest.mark.parametrize("msc_mode", [0, 1])
def test_batch_bitwise_equals_standalone(hip, msc_mode, persistent):
    from pyising_amd.batch import BatchScan

    specs = _mixed(hip)
    sims = [_make(s) for s in specs]
    refs = [_make(s) for s in specs]
    scan = BatchScan(sims, msc_mode=msc_mode, persistent=persistent)
    scan.sweep(3)
    scan.sweep(4)
    scan.synchronize()
    for sim, ref in zip(sims, refs, strict=True):
        ref.sweep(7)
        assert sim.t == ref.t == 7
        np.testing.assert_array_equal(sim.get_lattice(), ref.get_lattice())
    scan.close()


@pytest.mark.gpu
@pytest.mark.parametrize("case", ["int8-2d", "int8-3d-heatbath", "msc", "mixed"])
def test_dependency_ordered_split_lattices_match_oracle(hip, case):
    """Dependency-ordered launches where one lattice dominates the batch: it is cut into eight
    bands, one per XCD, whose edge tiles use the device-scope fences, while small lattices run
    whole on one XCD. Several launches (3 + 4 sweeps) and one long launch (25 sweeps) against
    the oracle, moments against the oracle's magnetisation after every sweep."""
    from pyising_amd.batch import BatchScan

    big = {"int8-2d": (hip.HipInt8, (512, 512), 1 / 2.269, "metropolis"),
           "int8-3d-heatbath": (hip.HipInt8, (64, 64, 64), 0.2216, "heatbath"),
           "msc": (hip.HipMultiSpin, (2048, 4096), 1 / 2.269, "metropolis"),
           "mixed": (hip.HipInt8, (1024, 1024), 1 / 2.3, "metropolis")}[case]
    specs = [(big[0], big[1], big[2], big[3], 900, O.random_lattice(big[1], 5))]
    if case == "mixed":
        specs += [s for s in _mixed(hip, seed0=300)[:8]]
    sims = [_make(s) for s in specs]
    scan = BatchScan(sims, persistent=True)
    refs = [_oracle(s) for s in specs]
    m_ref = [[] for _ in specs]
    for calls in ((3, 4), (25,)):
        for n in calls:
            scan.sweep(n)
            for i, r in enumerate(refs):
                for _ in range(n):
                    r.sweep(1)
                    m_ref[i].append(r.observables_exact()[0])
        scan.synchronize()
        for sim, ref in zip(sims, refs, strict=True):
            np.testing.assert_array_equal(sim.get_lattice(), ref.get_lattice())
    for mom, ms in zip(scan.moments(), m_ref, strict=True):
        m = np.array(ms, dtype=object)
        assert mom.samples == len(ms) == 32
        assert mom.sum_m == int(m.sum())
        assert mom.sum_abs_m == int(abs(m).sum())
        assert mom.sum_m2 == int((m * m).sum())
        assert mom.sum_m4 == int((m**4).sum())
    scan.close()


@pytest.mark.gpu
def test_batch_matches_oracle(hip):
    from pyising_amd.batch import BatchScan

    specs = _mixed(hip)[:8]
    sims = [_make(s) for s in specs]
    scan = BatchScan(sims, blocks_per_cu=2)
    scan.sweep(5)
    scan.synchronize()
    for spec, sim in zip(specs, sims, strict=True):
        cls, shape, beta, alg, seed, lat = spec
        ref = (O.MscOracle if cls is hip.HipMultiSpin else O.Int8Oracle)(
            shape, beta, seed=seed, algorithm=alg, init=lat)
        ref.sweep(5)
        np.testing.assert_array_equal(sim.get_lattice(), ref.get_lattice())
    scan.close()


@pytest.mark.gpu
@pytest.mark.parametrize("persistent", [False, True])
def test_batch_moments_are_exact(hip, persistent):
    from pyising_amd.batch import BatchScan

    specs = _mixed(hip)
    sims = [_make(s) for s in specs]
    scan = BatchScan(sims, persistent=persistent)
    refs = [_oracle(s) for s in specs]  # per-sweep exact M of the oracle
    n = 9
    scan.sweep(n)
    mom = scan.moments()
    assert scan.sweeps_done == n
    for ref, mo in zip(refs, mom, strict=True):
        ms = []
        for _ in range(n):
            ref.sweep(1)
            ms.append(ref.observables_exact()[0])
        ms = [int(m) for m in ms]
        assert mo.samples == n
        assert mo.sum_m == sum(ms)
        assert mo.sum_abs_m == sum(abs(m) for m in ms)
        assert mo.sum_m2 == sum(m * m for m in ms)
        assert mo.sum_m4 == sum(m**4 for m in ms)
    scan.reset_moments()
    assert all(m.samples == 0 for m in scan.moments())
    scan.close()


@pytest.mark.gpu
@pytest.mark.parametrize("persistent,moments", [(False, True), (True, True), (False, False),
                                                (True, False)])
def test_batch_graph_replays_continue_the_stream(hip, persistent, moments):
    from pyising_amd.batch import BatchScan

    specs = _mixed(hip)[:6]
    sims = [_make(s) for s in specs]
    refs = [_make(s) for s in specs]
    scan = BatchScan(sims, persistent=persistent, moments=moments)
    if not scan.capture(2):
        scan.close()
        pytest.skip("the runtime cannot capture the batch's launches into a graph")
    scan.replay(3)
    scan.sweep(1)  # a direct launch after replays continues the same counter
    scan.synchronize()
    for sim, ref in zip(sims, refs, strict=True):
        ref.sweep(7)
        np.testing.assert_array_equal(sim.get_lattice(), ref.get_lattice())
    assert scan.moments()[0].samples == (7 if moments else 0)
    assert scan.sweeps_done == 7
    scan.close()


@pytest.mark.gpu
def test_batch_rejects_bad_input(hip):
    from pyising_amd import _lib
    from pyising_amd.batch import BatchScan

    with pytest.raises(ValueError):
        BatchScan([])
    eager = hip.HipMultiSpinEager((128, 128), 0.4, seed=1)
    with pytest.raises(ValueError):
        BatchScan([eager])
    with pytest.raises(_lib.HipError):
        BatchScan([hip.HipInt8((16, 16), 0.4)], msc_cap=0)


# ------------------------------------------------------------------------ CPU (host logic) --
def test_struct_layout_matches_header():
    from pyising_amd.batch import IsingBatchLattice

    assert C.sizeof(IsingBatchLattice) == 96
    assert IsingBatchLattice.seed.offset == 24
    assert IsingBatchLattice.table.offset == 32
    assert IsingBatchLattice.c0.offset == 64
    assert IsingBatchLattice.t0.offset == 80
    assert IsingBatchLattice.vec.offset == 84
    assert IsingBatchLattice.max_sweeps.offset == 88


def test_moments_math():
    from pyising_amd.batch import Moments

    ms = np.array([4, -4, 2, -2, 4], dtype=np.int64)
    mo = Moments(samples=5, sum_m=int(ms.sum()), sum_abs_m=int(np.abs(ms).sum()),
                 sum_m2=int((ms**2).sum()), sum_m4=int((ms**4).sum()), n_spins=4)
    m = ms / 4.0
    assert mo.mean(1) == pytest.approx(m.mean())
    assert mo.abs_m == pytest.approx(np.abs(m).mean())
    assert mo.binder == pytest.approx(1 - (m**4).mean() / (3 * (m**2).mean() ** 2))
    ordered = Moments(samples=3, sum_m=3 * 16, sum_abs_m=3 * 16, sum_m2=3 * 256,
                      sum_m4=3 * 65536, n_spins=16)
    assert ordered.binder == pytest.approx(2 / 3)
    empty = Moments(0, 0, 0, 0, 0, 16)
    assert np.isnan(empty.binder)


@pytest.mark.gpu
@pytest.mark.parametrize("persistent", [False, True])
def test_sweep_limits_stop_each_lattice_at_its_own_count(hip, persistent):
    """Per-lattice sweep limits: every lattice ends bit-identical to the oracle after exactly
    min(limit, sweeps run) sweeps, and its moment sums cover exactly those sweeps."""
    from pyising_amd.batch import BatchScan

    specs = _mixed(hip)
    limits = [3, 0, 7, 1, 9, 2, 0, 5, 4, 6]
    sims = [_make(s) for s in specs]
    refs = [_make(s) for s in specs]
    scan = BatchScan(sims, persistent=persistent, limits=limits)
    scan.sweep(4)
    scan.sweep(4)
    scan.synchronize()
    mom = scan.moments()
    for sim, ref, lim, mo in zip(sims, refs, limits, mom, strict=True):
        n = 8 if lim == 0 else min(8, lim)
        ms = []
        for _ in range(n):
            ref.sweep(1)
            ms.append(int(ref.observables_exact()[0]))
        assert sim.t == ref.t == n
        np.testing.assert_array_equal(sim.get_lattice(), ref.get_lattice())
        assert mo.samples == n and mo.sum_m == sum(ms) and mo.sum_m2 == sum(m * m for m in ms)
    scan.close()


@pytest.mark.gpu
def test_reserved_batch_limits_across_assigns(hip):
    """A reserved batch re-assigned with fresh limits between steps (the serving pattern)."""
    from pyising_amd import _lib
    from pyising_amd.batch import ReservedBatch

    specs = _mixed(hip)
    sims = [_make(s) for s in specs]
    refs = [_make(s) for s in specs]
    st = _lib.Stream()
    rb = ReservedBatch(len(sims), 4096, st, msc_mode=1)
    left = [5, 12, 3, 8, 1, 10, 7, 2, 6, 4]
    while any(left):
        ids = [i for i, n in enumerate(left) if n]
        rb.assign([sims[i] for i in ids], [left[i] for i in ids])
        step = min(4, max(left[i] for i in ids))
        rb.sweep(step)
        for i in ids:
            left[i] -= min(step, left[i])
    st.sync()
    for sim, ref, n in zip(sims, refs, [5, 12, 3, 8, 1, 10, 7, 2, 6, 4], strict=True):
        ref.sweep(n)
        assert sim.t == ref.t == n
        np.testing.assert_array_equal(sim.get_lattice(), ref.get_lattice())
    rb.close()


@pytest.mark.gpu
@pytest.mark.parametrize("persistent", [False, True])
@pytest.mark.parametrize("k", [2, 4, 8])
def test_tiles_per_workgroup_change_nothing(hip, k, persistent):
    """k tiles per workgroup (the lattices-per-workgroup schedule; in the dependency-ordered
    form k consecutive tiles of one lattice share one wait): lattices and moment sums equal the
    one-tile form and the oracle, also with k changed between calls and under graph replay; bad
    k is rejected."""
    from pyising_amd import _lib
    from pyising_amd.batch import BatchScan

    specs = _mixed(hip)
    sims = [_make(s) for s in specs]
    base = [_make(s) for s in specs]
    refs = [_make(s) for s in specs]
    scan = BatchScan(sims, tiles_per_workgroup=k, persistent=persistent)
    ref_scan = BatchScan(base)
    scan.sweep(3)
    scan.set_tiles_per_workgroup(1)
    scan.sweep(2)
    scan.set_tiles_per_workgroup(k)
    assert scan.capture(2)
    scan.replay(2)
    ref_scan.sweep(9)
    scan.synchronize()
    ref_scan.synchronize()
    for sim, b, ref in zip(sims, base, refs, strict=True):
        ref.sweep(9)
        assert sim.t == b.t == ref.t == 9
        np.testing.assert_array_equal(sim.get_lattice(), ref.get_lattice())
        np.testing.assert_array_equal(b.get_lattice(), ref.get_lattice())
    for m, mb in zip(scan.moments(), ref_scan.moments(), strict=True):
        assert (m.samples, m.sum_m, m.sum_m2, m.sum_m4) == (mb.samples, mb.sum_m, mb.sum_m2, mb.sum_m4)
    lib = _lib.load_library()
    assert lib.ising_batch_set_tiles_per_workgroup(C.c_void_p(scan._h), 3) == -1
    scan.close()
    ref_scan.close()


@pytest.mark.gpu
def test_moment_sums_carry_past_64_bits(hip):
    """An ordered 16384^2 lattice has M^2 ~ 2^56, so 300 sweeps take sum M^2 past 2^64: the
    128-bit sums must equal sum M^2 and sum M^4 of the exact per-sweep M of the same lattice
    swept alone."""
    from pyising_amd.batch import BatchScan

    shape, beta, n = (16384, 16384), 1.0, 300
    sim = hip.HipMultiSpin(shape, beta, seed=5, init="cold")
    alone = hip.HipMultiSpin(shape, beta, seed=5, init="cold")
    scan = BatchScan([sim])
    scan.sweep(n)
    mo = scan.moments()[0]
    ms = []
    for _ in range(n):
        alone.sweep(1)
        ms.append(int(alone.observables_exact()[0]))
    assert mo.samples == n
    assert mo.sum_m2 == sum(m * m for m in ms) and mo.sum_m2 >= 2**64
    assert mo.sum_m4 == sum(m**4 for m in ms)
    assert mo.sum_abs_m == sum(abs(m) for m in ms) and mo.sum_m == sum(ms)
    scan.close()


def _random_mix(rng, hip):
    specs = []
    for i in range(int(rng.integers(5, 10))):
        kind = int(rng.integers(0, 3))
        beta = float(rng.uniform(0.1, 0.7))
        if kind == 0:
            shape = (2 * int(rng.integers(2, 40)), 8 * int(rng.integers(1, 40)))
            cls, alg = hip.HipInt8, ("metropolis", "heatbath")[int(rng.integers(0, 2))]
        elif kind == 1:
            shape = (2 * int(rng.integers(1, 5)), 2 * int(rng.integers(1, 8)),
                     8 * int(rng.integers(1, 4)))
            cls, alg, beta = hip.HipInt8, ("metropolis", "heatbath")[int(rng.integers(0, 2))], beta / 2
        else:
            shape = (2 * int(rng.integers(1, 60)), 128 * int(rng.integers(1, 40)))
            cls, alg = hip.HipMultiSpin, "metropolis"
        specs.append((cls, shape, beta, alg, int(rng.integers(0, 2**40)),
                      O.random_lattice(shape, int(rng.integers(0, 1000)))))
    return specs


@pytest.mark.gpu
@pytest.mark.parametrize("draw", range(6))
def test_random_batches_match_oracle(hip, draw):
    """Seeded random mixes: shapes, variants, sweep limits, tiles per workgroup, launch form and
    msc tile kind drawn at random; every lattice must equal the oracle after its own sweeps."""
    from pyising_amd.batch import BatchScan

    rng = np.random.default_rng(1000 + draw)
    specs = _random_mix(rng, hip)
    sims = [_make(s) for s in specs]
    refs = [_oracle(s) for s in specs]
    limits = [int(x) for x in rng.integers(0, 9, len(specs))]
    persistent = bool(draw % 3 == 2)
    scan = BatchScan(sims, msc_mode=int(rng.integers(0, 2)), persistent=persistent,
                     limits=limits, tiles_per_workgroup=int(rng.choice([1, 2, 4, 8])))
    scan.sweep(3)
    scan.sweep(4)
    scan.synchronize()
    for sim, ref, lim in zip(sims, refs, limits, strict=True):
        n = 7 if lim == 0 else min(7, lim)
        ref.sweep(n)
        assert sim.t == n
        assert sim.observables_exact() == tuple(ref.observables_exact())
        np.testing.assert_array_equal(sim.get_lattice(), ref.get_lattice())
    scan.close()


@pytest.mark.gpu
def test_members_read_after_the_batch_without_synchronize(hip):
    """Member reads straight after an asynchronous batched sweep (no synchronize) see the swept
    lattice: the member's stream waits for the batch's stream. A member's own sweep is refused
    while it belongs to the batch, and allowed again after close()."""
    from pyising_amd.batch import BatchScan

    specs = [(hip.HipMultiSpin, (2048, 2048), 1 / 2.269, "metropolis", 41,
              O.random_lattice((2048, 2048), 3)),
             (hip.HipInt8, (1024, 1024), 1 / 2.269, "heatbath", 42,
              O.random_lattice((1024, 1024), 4))]
    sims = [_make(s) for s in specs]
    refs = [_make(s) for s in specs]
    scan = BatchScan(sims)
    for _ in range(3):
        scan.sweep(20)
        for sim, ref in zip(sims, refs, strict=True):
            ref.sweep(20)
            np.testing.assert_array_equal(sim.get_lattice(), ref.get_lattice())
            assert sim.observables_exact() == ref.observables_exact()
    with pytest.raises(RuntimeError, match="belongs to a batch"):
        sims[0].sweep(1)
    scan.close()
    for sim, ref in zip(sims, refs, strict=True):
        sim.sweep(2)
        ref.sweep(2)
        np.testing.assert_array_equal(sim.get_lattice(), ref.get_lattice())
