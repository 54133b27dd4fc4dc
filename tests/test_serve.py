"""Serving loop: host logic (arena, admission, request stream) on the CPU; on the GPU, every
request's lattice and (M, E) served in a mix equal the same request run alone, and match the
oracle."""

import ctypes as C

import numpy as np
import pytest

from oracle import ising_oracle as O
from pyising_amd.serve import (
    ALIGN,
    AdmissionController,
    DeviceArena,
    Request,
    RequestRejected,
    poisson_requests,
)


def _req(rid, dim=2, L=64, variant="int8", algorithm="metropolis", T=2.3, sweeps=10, seed=None,
         arrival=0.0):
    return Request(rid=rid, dim=dim, L=L, variant=variant, algorithm=algorithm, T=T, sweeps=sweeps,
                   seed=1000 + rid if seed is None else seed, arrival=arrival)


# ------------------------------------------------------------------------------ arena (CPU) --
def test_arena_first_fit_and_coalescing():
    a = DeviceArena(10 * ALIGN)
    x = a.alloc(1)
    y = a.alloc(3 * ALIGN)
    z = a.alloc(2 * ALIGN + 1)
    assert (x, y, z) == (0, ALIGN, 4 * ALIGN)
    assert a.used == 7 * ALIGN and a.largest_free == 3 * ALIGN
    assert a.alloc(4 * ALIGN) is None
    a.release(y)
    a.check()
    assert a.alloc(4 * ALIGN) is None  # 3 + 3 granules free, not contiguous
    a.release(x)
    a.check()
    assert a.largest_free == 4 * ALIGN  # x and y coalesced
    assert a.alloc(4 * ALIGN) == 0
    a.release(z)
    a.release(0)
    a.check()
    assert a.used == 0 and a.largest_free == a.size
    with pytest.raises(KeyError):
        a.release(0)


def test_arena_random_churn_keeps_invariants():
    rng = np.random.default_rng(3)
    a = DeviceArena(1 << 20)
    live = {}
    for _ in range(3000):
        if live and rng.random() < 0.45:
            off = int(rng.choice(list(live)))
            a.release(off)
            del live[off]
        else:
            n = int(rng.integers(1, 40_000))
            off = a.alloc(n)
            if off is not None:
                assert off % ALIGN == 0 and off + n <= a.size
                for o, m in live.items():  # no overlap with live ranges
                    assert off + n <= o or o + m <= off
                live[off] = n
        a.check()
    for off in list(live):
        a.release(off)
    assert a.used == 0 and a._free == [[0, a.size]]


# -------------------------------------------------------------------------- admission (CPU) --
def test_footprint_is_the_aligned_state():
    from pyising_amd.backends.hip_backend import OBS_SCRATCH_BYTES as OBS

    r = _req(0, L=64)  # int8 2D: 2 colour arrays of 64*32 bytes, reduction scratch, 24 B result
    assert r.footprint() == 2 * 2048 + OBS + ALIGN
    m = _req(1, L=256, variant="msc")  # 2 x 256*2 words of 8 B
    assert m.footprint() == 2 * 4096 + OBS + ALIGN
    c = _req(2, dim=3, L=16)  # 2 x 16*16*8 bytes
    assert c.footprint() == 2 * 2048 + OBS + ALIGN


@pytest.mark.parametrize("bad", [
    dict(variant="fp16"), dict(algorithm="wolff"), dict(variant="msc", L=64),
    dict(variant="msc", dim=3, L=128), dict(variant="msc", L=128, algorithm="heatbath"),
    dict(L=30), dict(L=4), dict(sweeps=-1), dict(T=0.0), dict(dim=4),
])
def test_invalid_requests_are_rejected(bad):
    ctrl = AdmissionController(DeviceArena(1 << 24))
    with pytest.raises(RequestRejected):
        ctrl.submit(_req(0, **bad))
    assert not ctrl.pending


def test_oversize_and_duplicate_requests_are_rejected():
    ctrl = AdmissionController(DeviceArena(1 << 16), max_spins=1 << 14)
    with pytest.raises(RequestRejected, match="device memory"):
        ctrl.submit(_req(0, L=512))  # 256 KiB of state in a 64 KiB arena
    with pytest.raises(RequestRejected, match="at most"):
        ctrl.submit(_req(1, dim=3, L=32, variant="int8"))  # 32768 spins > 16384
    ctrl.submit(_req(2))
    with pytest.raises(RequestRejected, match="already"):
        ctrl.submit(_req(2))


def test_admission_by_memory_and_slots_fifo():
    big, small = _req(0, L=256).footprint(), _req(0, L=32).footprint()
    assert 3 * small < big
    arena = DeviceArena(2 * big)
    ctrl = AdmissionController(arena, max_slots=3, max_bypass=2)
    for i, L in enumerate([256, 256, 256, 32, 32, 32, 32]):
        ctrl.submit(_req(i, L=L))
    got = [r.rid for r, _ in ctrl.admit()]
    assert got == [0, 1]  # memory full: 2 x big; head (2) blocked, no room for small either
    assert arena.used == 2 * big
    ctrl.release(0)
    got = [r.rid for r, _ in ctrl.admit()]
    assert got == [2]  # FIFO: the head gets the freed memory first
    ctrl.release(1)
    got = [r.rid for r, _ in ctrl.admit()]
    assert got == [3, 4]  # memory for many small ones, but only 3 slots
    assert sorted(ctrl.active) == [2, 3, 4]
    assert arena.used == big + 2 * small
    assert [r.rid for r in ctrl.pending] == [5, 6]


def test_bypass_is_bounded():
    big, small = _req(0, L=256).footprint(), _req(0, L=32).footprint()
    arena = DeviceArena(big + small // 2)
    ctrl = AdmissionController(arena, max_slots=64, max_bypass=2)
    ctrl.submit(_req(0, L=32))
    assert [r.rid for r, _ in ctrl.admit()] == [0]
    ctrl.submit(_req(1, L=256))  # does not fit beside request 0
    for i in range(2, 8):
        ctrl.submit(_req(i, L=32))
    got = [r.rid for r, _ in ctrl.admit()]
    assert got == [2, 3]  # two small requests pass the blocked head, then the queue waits
    assert [r.rid for r in ctrl.pending][:1] == [1]
    for rid in (0, 2, 3):
        ctrl.release(rid)
    assert [r.rid for r, _ in ctrl.admit()][0] == 1  # the head goes first once it fits


def test_cancel_only_queued():
    ctrl = AdmissionController(DeviceArena(1 << 20), max_slots=1)
    ctrl.submit(_req(0))
    ctrl.submit(_req(1))
    ctrl.admit()
    assert not ctrl.cancel(0)  # running
    assert ctrl.cancel(1)
    assert not ctrl.cancel(1)
    assert not ctrl.pending


def test_poisson_stream_follows_the_mix():
    reqs = poisson_requests(4000, rate=200.0, seed=5)
    assert reqs == poisson_requests(4000, rate=200.0, seed=5)
    gaps = np.diff([0.0] + [r.arrival for r in reqs])
    assert abs(gaps.mean() - 1 / 200) < 0.1 / 200
    two = [r for r in reqs if r.dim == 2]
    three = [r for r in reqs if r.dim == 3]
    assert abs(len(two) / len(reqs) - 0.7) < 0.03
    assert {r.L for r in two} == {2**k for k in range(5, 13)}
    assert {r.L for r in three} == {16, 32, 64, 128, 256}
    assert all(r.L % 128 == 0 for r in two if r.variant == "msc")
    assert all(r.variant == "int8" for r in two if r.L < 128)
    assert {r.algorithm for r in three} == {"metropolis", "heatbath"}
    assert all(r.algorithm == "metropolis" for r in two)
    assert all(100 <= r.sweeps <= 10_000 for r in reqs)
    for r in reqs:
        r.validate()


def test_reserved_batch_capacity_bound():
    """The server reserves arena/2048 + slots tiles: check that bound against the tile counts
    the C library computes for the largest lattices of each kind (host-only entry point)."""
    from pyising_amd import _lib
    from pyising_amd.batch import IsingBatchLattice

    lib = _lib.load_library()
    cases = [(0, 2, 1, 4096, 4096), (0, 3, 256, 256, 256), (1, 2, 1, 4096, 4096),
             (0, 2, 1, 32, 32), (0, 2, 1, 8, 24), (1, 2, 1, 2, 128)]
    for kind, ndim, lz, ly, lx in cases:
        d = IsingBatchLattice()
        d.kind, d.ndim, d.lz, d.ly, d.lx = kind, ndim, lz, ly, lx
        d.c0 = d.c1 = 256
        tiles = C.c_longlong()
        assert lib.ising_batch_tiles(C.byref(d), 1, 0, C.byref(tiles)) == 0
        n = lz * ly * lx
        state = 2 * (n // 16 if kind == 1 else n // 2)
        assert tiles.value <= state // 2048 + 1, (kind, ndim, lz, ly, lx, tiles.value)
    d = IsingBatchLattice()
    d.kind, d.ndim, d.lz, d.ly, d.lx = 1, 2, 1, 64, 96  # msc needs lx % 128 == 0
    d.c0 = d.c1 = 256
    assert lib.ising_batch_tiles(C.byref(d), 1, 0, C.byref(C.c_longlong())) == -1


# ------------------------------------------------------------------------------------ GPU --
def _mix():
    """Every kind of lattice, uneven sizes and sweep counts (some not multiples of the step)."""
    specs = [
        dict(dim=2, L=64, variant="int8", algorithm="metropolis", T=2.2, sweeps=23),
        dict(dim=2, L=32, variant="int8", algorithm="metropolis", T=3.0, sweeps=5),
        dict(dim=2, L=128, variant="msc", algorithm="metropolis", T=2.269, sweeps=31),
        dict(dim=2, L=256, variant="msc", algorithm="metropolis", T=1.9, sweeps=17),
        dict(dim=3, L=16, variant="int8", algorithm="heatbath", T=4.5, sweeps=12),
        dict(dim=3, L=16, variant="int8", algorithm="metropolis", T=4.0, sweeps=40),
        dict(dim=2, L=256, variant="int8", algorithm="heatbath", T=2.4, sweeps=9),
        dict(dim=3, L=32, variant="int8", algorithm="metropolis", T=5.0, sweeps=6),
        dict(dim=2, L=512, variant="msc", algorithm="metropolis", T=2.5, sweeps=14),
        dict(dim=2, L=96, variant="int8", algorithm="metropolis", T=2.0, sweeps=0),
        dict(dim=2, L=128, variant="int8", algorithm="metropolis", T=2.269, sweeps=1),
        dict(dim=3, L=8, variant="int8", algorithm="heatbath", T=3.5, sweeps=27),
    ]
    return [Request(rid=i, seed=7 + 31 * i, arrival=0.002 * i, **s) for i, s in enumerate(specs)]


def _check_against_alone(rep, reqs):
    from pyising_amd.serve import run_alone

    assert sorted(r.rid for r in rep.results) == [r.rid for r in reqs]
    by = {r.rid: r for r in rep.results}
    for q in reqs:
        M, E, lat = run_alone(q, lattice=True)
        got = by[q.rid]
        assert (got.M, got.E) == (M, E), q
        np.testing.assert_array_equal(got.lattice, lat, err_msg=str(q))


@pytest.mark.gpu
@pytest.mark.parametrize("step_sweeps,max_slots", [(7, 64), (4, 3), (100, 5)])
def test_mix_equals_alone(step_sweeps, max_slots):
    from pyising_amd.serve import Server

    reqs = _mix()
    srv = Server(budget_bytes=64 << 20, max_slots=max_slots, step_sweeps=step_sweeps,
                 max_spins=1 << 20)
    rep = srv.run(reqs, realtime=False, keep_lattices=True, timeout=120)
    srv.close()
    assert rep.max_active <= max_slots
    _check_against_alone(rep, reqs)


@pytest.mark.gpu
def test_memory_bound_admission_and_oracle():
    from pyising_amd.serve import Server

    reqs = _mix()
    budget = max(r.footprint() for r in reqs) + 2 * ALIGN * 4  # about one lattice at a time
    srv = Server(budget_bytes=budget, max_slots=64, step_sweeps=5, max_spins=1 << 20)
    rep = srv.run(reqs, realtime=False, keep_lattices=True, timeout=120)
    srv.close()
    assert srv.arena.used == 0
    by = {r.rid: r for r in rep.results}
    for q in reqs:
        cls = O.MscOracle if q.variant == "msc" else O.Int8Oracle
        ref = cls(q.shape, 1.0 / q.T, seed=q.seed, algorithm=q.algorithm, init="hot")
        ref.sweep(q.sweeps)
        np.testing.assert_array_equal(by[q.rid].lattice, ref.get_lattice(), err_msg=str(q))
        s = by[q.rid].lattice.astype(np.int64)
        assert by[q.rid].M == int(s.sum())


@pytest.mark.gpu
def test_realtime_arrivals_join_running_steps():
    from pyising_amd.serve import Server

    reqs = _mix()
    for i, r in enumerate(reqs):
        r.arrival = 0.004 * i
        r.sweeps = r.sweeps * 5
    # a long request first, so that the later arrivals meet running steps
    reqs[0] = Request(rid=0, dim=2, L=256, variant="int8", algorithm="metropolis", T=2.269,
                      sweeps=4000, seed=99, arrival=0.0)
    srv = Server(budget_bytes=64 << 20, max_slots=64, step_sweeps=10, max_spins=1 << 20)
    rep = srv.run(reqs, realtime=True, keep_lattices=True, timeout=120)
    srv.close()
    assert rep.max_active >= 2
    assert rep.admitted_during_steps > 0
    assert all(r.finished >= r.admitted >= r.arrival for r in rep.results)
    _check_against_alone(rep, reqs)


@pytest.mark.gpu
def test_server_rejects_without_stopping():
    from pyising_amd.serve import Server

    reqs = _mix()[:3] + [_req(50, L=4096), _req(51, L=30)]
    srv = Server(budget_bytes=8 << 20, max_slots=8, step_sweeps=5, max_spins=1 << 20)
    rep = srv.run(reqs, realtime=False, timeout=60)
    srv.close()
    assert sorted(rep.rejected) == [50, 51]
    assert sorted(r.rid for r in rep.results) == [0, 1, 2]


def test_gpu_mix_is_valid():
    for r in _mix():
        r.validate()


def test_batched_lattices_are_bounded_for_exact_moments():
    """A batched lattice holds at most 2^28 spins (16384^2): the 128-bit M^4 sum then stays exact
    for 65536 sweeps. Host-only entry point, no GPU needed."""
    import ctypes as C

    from pyising_amd import _lib
    from pyising_amd.batch import IsingBatchLattice

    lib = _lib.load_library()
    tiles = C.c_longlong()
    for (lx, ok) in ((16384, True), (16512, False)):
        d = IsingBatchLattice()
        d.kind, d.ndim, d.lz, d.ly, d.lx = 1, 2, 1, 16384, lx
        d.c0 = d.c1 = 256
        rc = lib.ising_batch_tiles(C.byref(d), 1, 0, C.byref(tiles))
        assert (rc == 0) == ok, (lx, rc)
        if not ok:
            assert rc == -3 and b"2^28" in lib.ising_last_error()
