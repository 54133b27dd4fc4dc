"""Batched temperature scans: many lattices of mixed size, dimension, variant and temperature
swept together on the GPU (``csrc/batch_kernels.hip``): one launch per half sweep, or
dependency-ordered launches of many sweeps.

No counterpart in the reference, which sweeps one lattice per launch sequence
(src/pyising/backends/base.py:85-90). Every lattice keeps its own seed and sweep counter, so a
batched lattice is bit-identical to the same lattice swept alone; Binder sums (M, |M|, M^2 and a
128-bit M^4) are accumulated on the device after every sweep.

>>> scan = BatchScan([HipInt8((64, 64), 1 / T, seed=i) for i, T in enumerate(temps)])
>>> scan.sweep(10_000)
>>> scan.moments()[0]["binder"]
"""

from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _lib
from .backends.base import IsingSimulation
from .backends.hip_backend import HipInt8, HipMultiSpin


class IsingBatchLattice(C.Structure):
    """Mirror of ``IsingBatchLattice`` in include/ising_hip.h."""

    _fields_ = [("kind", C.c_int), ("ndim", C.c_int), ("heatbath", C.c_int), ("lz", C.c_int),
                ("ly", C.c_int), ("lx", C.c_int), ("seed", C.c_ulonglong),
                ("table", C.c_uint32 * 7), ("c0", C.c_void_p), ("c1", C.c_void_p),
                ("t0", C.c_uint32), ("vec", C.c_int), ("max_sweeps", C.c_uint32)]


#: msc tile kinds of IsingBatchLattice.kind (1 = the batch's default tile)
MSC_TILE_KIND = {"default": 1, "lockstep": 2, "queue": 3}


def _describe(sim: IsingSimulation, msc_tile: str = "default", vec: int = 0,
              max_sweeps: int = 0) -> IsingBatchLattice:
    d = IsingBatchLattice()
    shape = sim.shape
    d.ndim = len(shape)
    d.lz = shape[0] if d.ndim == 3 else 1
    d.ly, d.lx = shape[-2], shape[-1]
    d.seed = sim.seed
    for i, v in enumerate(sim.table):
        d.table[i] = int(v)
    if isinstance(sim, HipMultiSpin):
        if sim.mode == 2:
            raise ValueError("the eager ablation is not batched")
        d.kind, d.heatbath = MSC_TILE_KIND[msc_tile], 0
        d.c0, d.c1 = sim.w[0].ptr, sim.w[1].ptr
    elif isinstance(sim, HipInt8):
        d.kind = 0
        d.vec = int(vec)
        d.heatbath = int(sim.algorithm == "heatbath")
        d.c0, d.c1 = sim.c[0].ptr, sim.c[1].ptr
    else:
        raise TypeError(f"{type(sim).__name__} is not a hip simulation")
    d.t0 = sim.t & 0xFFFFFFFF
    if not 0 <= int(max_sweeps) <= 0xFFFFFFFF:
        raise ValueError("max_sweeps must be in [0, 2^32)")
    d.max_sweeps = int(max_sweeps)
    return d


@dataclass
class Moments:
    """Per-lattice sums over sweeps (exact integers) and derived per-spin moments."""

    samples: int
    sum_m: int
    sum_abs_m: int
    sum_m2: int
    sum_m4: int
    n_spins: int

    def mean(self, k: int) -> float:
        s = {1: self.sum_m, 2: self.sum_m2, 4: self.sum_m4}[k]
        return s / self.samples / self.n_spins**k if self.samples else float("nan")

    @property
    def abs_m(self) -> float:
        return self.sum_abs_m / self.samples / self.n_spins if self.samples else float("nan")

    @property
    def binder(self) -> float:
        """U_4 = 1 - <m^4> / (3 <m^2>^2) (src/pyising/observables.py:39-42)."""
        m2, m4 = self.mean(2), self.mean(4)
        return 1.0 - m4 / (3.0 * m2 * m2) if self.samples and m2 > 0 else float("nan")

    def as_dict(self) -> dict:
        return {"samples": self.samples, "m": self.mean(1), "abs_m": self.abs_m,
                "m2": self.mean(2), "m4": self.mean(4), "binder": self.binder}


class _Members:
    """Membership of simulations in a batch that sweeps them on its own stream.

    While a simulation belongs to a batch, its own sweeps are refused (the batch holds its
    sweep counter) and its other device work (reading or writing the lattice, observables) is
    queued behind the batch's last recorded sweep: the batch's stream is a non-blocking stream,
    which the simulation's stream would otherwise not wait for."""

    _done = None
    _unchecked = False  # dependency-ordered launches queued since the last status check
    sims: list = []

    def _mark(self) -> None:
        if self._done is None:
            self._done = _lib.Event()
        self._done.record(self.stream.ptr)

    def order_after(self, stream_ptr) -> None:
        """Work queued on ``stream_ptr`` from now on waits for the batch's last recorded
        sweep (on the device; the host does not wait, except to check dependency-ordered
        launches that have not been checked yet)."""
        if self._unchecked:
            self.check()
        if self._done is not None:
            _lib.check(self._L.ising_stream_wait_event(stream_ptr, self._done.ptr),
                       "stream_wait_event")

    def _join(self, sims) -> None:
        for s in sims:
            if getattr(s, "_batch", None) not in (None, self):
                raise ValueError("a lattice can belong to one batch at a time")
        for s in sims:
            s._batch = self

    def _leave(self, sims) -> None:
        for s in sims:
            if getattr(s, "_batch", None) is self:
                s._batch = None
                self.order_after(s.stream_ptr)

    def _release_members(self) -> None:
        self._leave(self.sims)
        if self._done is not None:
            self._done.close()
            self._done = None


class BatchScan(_Members):
    """Sweep a list of hip simulations together, every lattice in the same launches.

    ``blocks_per_cu``: accepted for compatibility; the launch sizes follow the tile counts.
    ``msc_mode``: 0 = queue compaction (512-word tiles), 1 = lockstep tiles.
    ``persistent``: True = dependency-ordered launches (up to 512 sweeps per launch: each tile
    waits only for the tiles of its own lattice holding its neighbour rows in the previous half
    sweep, through per-tile completion counters); False (default) = one launch per half sweep
    over all lattices plus one accumulation launch per sweep.
    The simulations keep ownership of their device arrays; while they belong to the batch (until
    ``close``) their reads and writes (``get_lattice``, ``observables``, ``set_lattice``) are
    queued behind the batch's last sweep, their own ``sweep`` is refused, and their ``t``
    advances with every batched sweep. Thresholds are taken at construction: ``set_beta`` on a
    member afterwards does not reach the batch.
    """

    def __init__(self, sims, blocks_per_cu: int | None = None, msc_mode: int = 0,
                 msc_cap: int = 4, persistent: bool = False, stream: _lib.Stream | None = None,
                 msc_tiles: dict | None = None, int8_vecs: dict | None = None,
                 moments: bool = True, limits=None, tiles_per_workgroup: int = 1):
        """``msc_tiles``: optional {shape: "queue" | "lockstep"} choosing the tile of the msc
        lattices of that shape (others use ``msc_mode``); ``int8_vecs``: optional {shape: 16 |
        8 | 4} sites per lane of the int8 lattices of that shape (others: the widest width that
        fills a tile); e.g. from :func:`pyising_amd.autotune.tune_batch`. ``moments``: False
        skips the per-sweep accumulation of M moments (one launch per sweep fewer; moments()
        then reports no samples). ``limits``: optional per-lattice number of sweeps it takes
        part in (0 = no limit); past it a lattice and its moment sums stay as they are.
        ``tiles_per_workgroup``: launch form, tiles each workgroup runs in sequence (1, 2, 4,
        8; the autotuner's lattices-per-workgroup schedule). No setting changes a result."""
        self.sims = list(sims)
        if not self.sims:
            raise ValueError("empty batch")
        self._L = _lib.lib()
        tiles, vecs = msc_tiles or {}, int8_vecs or {}
        lim = [0] * len(self.sims) if limits is None else [int(x) for x in limits]
        if len(lim) != len(self.sims):
            raise ValueError("one sweep limit per lattice")
        arr = (IsingBatchLattice * len(self.sims))(
            *[_describe(s, tiles.get(tuple(s.shape), "default"), vecs.get(tuple(s.shape), 0), n)
              for s, n in zip(self.sims, lim, strict=True)])
        self._left = [n if n else None for n in lim]
        h = C.c_void_p()
        _lib.check(self._L.ising_batch_create(arr, len(self.sims), int(blocks_per_cu or 0),
                                              int(msc_mode), int(msc_cap), C.byref(h)),
                   "batch_create")
        self._h = h.value
        _lib.check(self._L.ising_batch_set_persistent(self._h, int(bool(persistent))),
                   "batch_set_persistent")
        _lib.check(self._L.ising_batch_set_moments(self._h, int(bool(moments))),
                   "batch_set_moments")
        if int(tiles_per_workgroup) != 1:
            _lib.check(self._L.ising_batch_set_tiles_per_workgroup(self._h,
                                                                   int(tiles_per_workgroup)),
                       "batch_set_tiles_per_workgroup")
        self.tiles_per_workgroup = int(tiles_per_workgroup)
        self.persistent = bool(persistent)
        self.n_spins = np.array([s.n_spins for s in self.sims], dtype=np.int64)
        self._own_stream = stream is None
        self.stream = stream if stream is not None else _lib.Stream()
        self.graph_sweeps = 0
        self._join(self.sims)

    @classmethod
    def tuned(cls, sims, sweeps: int = 50, use_cache: bool = True) -> "BatchScan":
        """A scan with the autotuner's choices for this batch (see pyising_amd.autotune); the
        graph, if any, is captured already (``graph_sweeps`` > 0: use :meth:`replay`)."""
        from .autotune import tune_batch

        t = tune_batch(sims, sweeps=sweeps, use_cache=use_cache)
        scan = cls(sims, persistent=t.persistent, msc_tiles=t.msc_tiles, int8_vecs=t.int8_vecs,
                   tiles_per_workgroup=t.tiles_per_workgroup)
        scan.tuning = t
        if t.graph_sweeps and not scan.capture(t.graph_sweeps):
            scan.graph_sweeps = 0
        return scan

    def info(self) -> dict:
        g, t, m = C.c_int(), C.c_int(), C.c_int()
        _lib.check(self._L.ising_batch_info(self._h, C.byref(g), C.byref(t), C.byref(m)))
        return {"grid": g.value, "tiles_per_half_sweep": t.value, "max_blocks_per_cu": m.value}

    def set_tiles_per_workgroup(self, k: int) -> None:
        """Tiles each workgroup of the launch form runs in sequence (1, 2, 4 or 8). Waits for
        the device; a captured graph is dropped (capture again to replay)."""
        _lib.check(self._L.ising_batch_set_tiles_per_workgroup(self._h, int(k)),
                   "batch_set_tiles_per_workgroup")
        self.tiles_per_workgroup = int(k)
        self.graph_sweeps = 0

    def _advance(self, n: int) -> None:
        for i, s in enumerate(self.sims):
            left = self._left[i]
            step = n if left is None else min(n, left)
            s.t += step
            if left is not None:
                self._left[i] = left - step

    def sweep(self, n: int = 1) -> None:
        """n sweeps of every lattice (asynchronous on the batch stream)."""
        n = int(n)
        if n <= 0:
            return
        _lib.check(self._L.ising_batch_sweeps(self._h, n, self.stream.ptr), "batch_sweeps")
        self._unchecked = self.persistent
        self._mark()
        self._advance(n)

    def _status(self) -> int:
        failed = C.c_int()
        _lib.check(self._L.ising_batch_status(self._h, C.byref(failed), self.stream.ptr),
                   "batch_status")
        self._unchecked = False
        return failed.value

    def check(self) -> None:
        """Raise if a dependency-ordered launch failed (a bounded wait gave up, or a tile did
        not publish its last half sweep); the lattices and moments are then invalid.
        Synchronous."""
        failed = self._status()
        if failed:
            raise _lib.HipError(f"dependency-ordered batch launch failed (status {failed}): "
                                "the lattices are invalid; use persistent=False")

    def capture(self, sweeps_per_replay: int) -> bool:
        """Capture one launch into a hipGraph; False if the runtime cannot capture it."""
        rc = self._L.ising_batch_graph(self._h, int(sweeps_per_replay), self.stream.ptr)
        if rc < 0:
            self.graph_sweeps = 0
            return False
        self.graph_sweeps = int(sweeps_per_replay)
        return True

    def replay(self, replays: int = 1) -> None:
        if not self.graph_sweeps:
            raise RuntimeError("no captured graph: call capture() first")
        _lib.check(self._L.ising_batch_replay(self._h, int(replays), self.stream.ptr),
                   "batch_replay")
        self._unchecked = self.persistent
        self._mark()
        self._advance(self.graph_sweeps * int(replays))

    def synchronize(self) -> None:
        self.stream.sync()
        if self._unchecked:
            self.check()

    def moments(self) -> list[Moments]:
        if self._unchecked:
            self.check()
        out = np.zeros((len(self.sims), 7), dtype=np.int64)
        done = C.c_longlong()
        _lib.check(self._L.ising_batch_stats(self._h, out.ctypes.data, C.byref(done),
                                             self.stream.ptr), "batch_stats")
        res = []
        u = out.view(np.uint64)
        for i, row in enumerate(out):
            m2 = int(u[i, 2]) | (int(u[i, 3]) << 64)  # 128-bit sums as (lo, hi)
            m4 = int(u[i, 4]) | (int(u[i, 5]) << 64)
            res.append(Moments(samples=int(row[6]), sum_m=int(row[0]), sum_abs_m=int(u[i, 1]),
                               sum_m2=m2, sum_m4=m4, n_spins=int(self.n_spins[i])))
        self.sweeps_done = done.value
        return res

    def reset_moments(self) -> None:
        _lib.check(self._L.ising_batch_reset_stats(self._h, self.stream.ptr), "batch_reset")

    def close(self) -> None:
        """Wait for the batch, release its device tables (and its stream, if it made one);
        the member simulations leave the batch and may sweep on their own again."""
        if getattr(self, "_h", None):
            self.stream.sync()
            failed = self._status() if self._unchecked else 0
            self._L.ising_batch_destroy(self._h)
            self._h = None
            self._release_members()
            if getattr(self, "_own_stream", False):
                self.stream.close()
            if failed:
                raise _lib.HipError(f"dependency-ordered batch launch failed (status {failed})")

    def __del__(self):
        if not _lib.releasing_allowed():
            return
        try:
            self.close()
        except Exception:  # noqa: BLE001 - a failing release in a finalizer is dropped
            pass


def _describe_all(sims, limits=None) -> C.Array:
    limits = [0] * len(sims) if limits is None else [0 if n is None else int(n) for n in limits]
    if len(limits) != len(sims):
        raise ValueError("one sweep limit per lattice")
    return (IsingBatchLattice * max(1, len(sims)))(
        *[_describe(s, max_sweeps=n) for s, n in zip(sims, limits, strict=True)])


def batch_tiles(sims, msc_mode: int = 0) -> int:
    """Tiles of one half sweep of these simulations (host computation, no device call)."""
    out = C.c_longlong()
    _lib.check(_lib.load_library().ising_batch_tiles(_describe_all(sims), len(sims),
                                                     int(msc_mode), C.byref(out)), "batch_tiles")
    return out.value


class ReservedBatch(_Members):
    """A batch with fixed device capacity whose lattice list is replaced between sweeps
    (continuous batching). ``assign`` and ``sweep`` are stream-ordered on ``stream`` and allocate
    nothing; a sweep queued before an ``assign`` still runs the old list.
    """

    def __init__(self, max_lattices: int, max_tiles: int, stream: _lib.Stream, msc_mode: int = 0,
                 msc_cap: int = 4, blocks_per_cu: int | None = None, moments: bool = False):
        """``moments``: accumulate M moments after every sweep (off by default: a serving loop
        reads each lattice's own exact (M, E) when it finishes)."""
        self._L = _lib.lib()
        self.stream = stream
        self.msc_mode = int(msc_mode)
        h = C.c_void_p()
        _lib.check(self._L.ising_batch_reserve(int(max_lattices), int(max_tiles),
                                               int(blocks_per_cu or 0), self.msc_mode,
                                               int(msc_cap), C.byref(h)), "batch_reserve")
        self._h = h.value
        _lib.check(self._L.ising_batch_set_moments(self._h, int(bool(moments))),
                   "batch_set_moments")
        self.max_lattices, self.max_tiles = int(max_lattices), int(max_tiles)
        self.sims: list[IsingSimulation] = []
        self._left: list[int | None] = []  # sweeps each lattice may still take (None: any)

    def assign(self, sims, limits=None) -> None:
        """Replace the lattice list. ``limits``: per lattice, the number of sweeps it takes
        part in from now on (0 or None: no limit); a lattice past its limit sits the batch's
        further sweeps out, so one batch sweep count can serve requests of any length."""
        sims = list(sims)
        desc = _describe_all(sims, limits)
        _lib.check(self._L.ising_batch_assign(self._h, desc, len(sims), self.stream.ptr),
                   "batch_assign")
        self._mark()
        self._leave([s for s in self.sims if all(s is not t for t in sims)])
        self._join(sims)
        self.sims = sims
        self._left = [int(n) if n else None for n in (limits or [0] * len(sims))]

    def sweep(self, n: int = 1) -> None:
        n = int(n)
        if n <= 0 or not self.sims:
            return
        _lib.check(self._L.ising_batch_sweeps(self._h, n, self.stream.ptr), "batch_sweeps")
        self._mark()
        for i, s in enumerate(self.sims):
            left = self._left[i]
            if left is None:
                s.t += n
            else:
                s.t += min(n, left)
                self._left[i] = left - min(n, left)

    def info(self) -> dict:
        g, t, m = C.c_int(), C.c_int(), C.c_int()
        _lib.check(self._L.ising_batch_info(self._h, C.byref(g), C.byref(t), C.byref(m)))
        return {"grid": g.value, "tiles_per_half_sweep": t.value, "max_blocks_per_cu": m.value}

    def close(self) -> None:
        if getattr(self, "_h", None):
            self.stream.sync()
            self._L.ising_batch_destroy(self._h)
            self._h = None
            self._release_members()

    def __del__(self):
        if not _lib.releasing_allowed():
            return
        try:
            self.close()
        except Exception:  # noqa: BLE001 - a failing release in a finalizer is dropped
            pass


def temperature_scan_lattices(int8_sizes=(64, 128, 256, 512, 1024, 2048), msc_sizes=(1024, 4096),
                              temps=None, base_seed: int = 1000, init: str = "hot"):
    """The lattices of a Binder temperature scan: 2D int8 lattices of every size in
    ``int8_sizes`` and 2D msc lattices of every size in ``msc_sizes``, each at every temperature
    (default 16 uniformly spaced in [2.20, 2.34]), seed = base_seed + index."""
    if temps is None:
        temps = np.linspace(2.20, 2.34, 16)
    sims = []
    for L in int8_sizes:
        for T in temps:
            sims.append(HipInt8((L, L), 1.0 / T, seed=base_seed + len(sims), init=init))
    for L in msc_sizes:
        for T in temps:
            sims.append(HipMultiSpin((L, L), 1.0 / T, seed=base_seed + len(sims), init=init))
    return sims
