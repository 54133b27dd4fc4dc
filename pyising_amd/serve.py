"""Continuous-batching serving loop for Ising simulation requests (no reference counterpart; the
reference runs one simulation per call, src/pyising/backends/base.py:85-90).

A request asks for ``sweeps`` sweeps of one lattice (2D or 3D, int8 or msc, a temperature, a
seed, a hot start) and gets back the exact final (M, E). Design:

* Memory. The server makes one device allocation for request state (``budget_bytes``) and one
  staging buffer (hot starts are generated there on the device) at start-up, and nothing after
  that: a host-side first-fit arena (:class:`DeviceArena`) hands every admitted request one
  region for its colour arrays and reduction scratch, returned when the request finishes.
  Admission is by that footprint, so a request is admitted exactly when its state fits the
  free memory and a batch slot is free.
* Streams. Admissions (device hot start, split into colour arrays) and result reads run on a
  setup stream; batched sweeps run on a work stream. Both are non-blocking streams, and the loop
  calls no hipMalloc/hipFree, so admission work overlaps the sweeps on the device.
* Continuous batching. Every step runs ``step_sweeps`` batch sweeps (fewer if every active
  request finishes sooner) through the batched kernels of :mod:`pyising_amd.batch`, one launch
  per half sweep over all lattices; each lattice carries its own sweep limit, so a request that
  needs fewer sweeps than the step sits the rest of the step out instead of cutting the step
  short for all the others. While a step runs, new arrivals are admitted and prepared;
  they join at the next step boundary, when the batch's lattice table is replaced in place
  (``ising_batch_assign``, stream-ordered, no allocation). Finished requests are read and freed
  between steps.
* Exactness. Every lattice keeps its own seed and sweep counter, so a request's lattice and
  (M, E) are bit-identical whether it ran alone or in any mix (``run_alone`` is the check).
"""

from __future__ import annotations

import math
import time
from collections import deque
from dataclasses import dataclass, field

import numpy as np

from .lattice import BETA_C_3D, TC_2D, check_shape

ALIGN = 256  # bytes: every device sub-allocation starts on this boundary


def _align(n: int) -> int:
    return (int(n) + ALIGN - 1) // ALIGN * ALIGN


class RequestRejected(ValueError):
    """The request is invalid or can never fit the server."""


@dataclass
class Request:
    rid: int
    dim: int
    L: int
    variant: str        # "int8" or "msc"
    algorithm: str      # "metropolis" or "heatbath"
    T: float
    sweeps: int
    seed: int
    arrival: float = 0.0  # seconds after the stream starts

    @property
    def shape(self) -> tuple[int, ...]:
        return (self.L,) * self.dim

    @property
    def n_spins(self) -> int:
        return self.L**self.dim

    def allocations(self) -> list[int]:
        """The device allocations of its simulation: the reduction scratch, the (M, E) result
        and the two colour arrays (int8: N/2 bytes each; msc: N/16 bytes each)."""
        from .backends.hip_backend import OBS_OUT_BYTES, OBS_SCRATCH_BYTES

        colour = self.n_spins // 16 if self.variant == "msc" else self.n_spins // 2
        return [OBS_SCRATCH_BYTES, OBS_OUT_BYTES, colour, colour]

    def footprint(self) -> int:
        """Device bytes held while admitted (every allocation 256-byte aligned)."""
        return sum(_align(b) for b in self.allocations())

    def validate(self) -> None:
        if self.variant not in ("int8", "msc"):
            raise RequestRejected(f"unknown variant {self.variant!r}")
        if self.algorithm not in ("metropolis", "heatbath"):
            raise RequestRejected(f"unknown algorithm {self.algorithm!r}")
        if self.variant == "msc" and (self.dim != 2 or self.algorithm != "metropolis"):
            raise RequestRejected("msc supports 2D Metropolis only")
        if self.sweeps < 0 or not self.T > 0:
            raise RequestRejected("sweeps must be >= 0 and T > 0")
        try:
            check_shape(self.shape, 128 if self.variant == "msc" else 8)
        except ValueError as exc:
            raise RequestRejected(str(exc)) from exc


@dataclass
class Result:
    rid: int
    M: int
    E: int
    n_spins: int
    sweeps: int
    arrival: float
    admitted: float
    finished: float
    lattice: np.ndarray | None = None

    @property
    def latency(self) -> float:
        return self.finished - self.arrival

    @property
    def m(self) -> float:
        return self.M / self.n_spins

    @property
    def e(self) -> float:
        return self.E / self.n_spins


class DeviceArena:
    """First-fit allocator of byte ranges in [0, size), 256-byte granules, with coalescing
    frees (host bookkeeping only: the server maps offsets onto one device allocation)."""

    def __init__(self, size: int):
        self.size = int(size) // ALIGN * ALIGN
        if self.size <= 0:
            raise ValueError("arena size must be at least 256 bytes")
        self._free: list[list[int]] = [[0, self.size]]  # [offset, length], sorted by offset
        self._live: dict[int, int] = {}

    @property
    def used(self) -> int:
        return sum(self._live.values())

    @property
    def largest_free(self) -> int:
        return max((n for _, n in self._free), default=0)

    def alloc(self, nbytes: int) -> int | None:
        n = _align(max(1, nbytes))
        for i, (off, length) in enumerate(self._free):
            if length >= n:
                if length == n:
                    del self._free[i]
                else:
                    self._free[i] = [off + n, length - n]
                self._live[off] = n
                return off
        return None

    def release(self, off: int) -> None:
        n = self._live.pop(off)  # KeyError on a double free
        i = 0
        while i < len(self._free) and self._free[i][0] < off:
            i += 1
        self._free.insert(i, [off, n])
        if i + 1 < len(self._free) and off + n == self._free[i + 1][0]:
            self._free[i][1] += self._free.pop(i + 1)[1]
        if i > 0 and self._free[i - 1][0] + self._free[i - 1][1] == off:
            self._free[i - 1][1] += self._free.pop(i)[1]

    def check(self) -> None:
        """Invariants: free blocks sorted, disjoint, coalesced; free + live = size."""
        prev_end = -1
        for off, n in self._free:
            assert n > 0 and off > prev_end, "free list unsorted, overlapping or not coalesced"
            prev_end = off + n
        assert sum(n for _, n in self._free) + self.used == self.size


class AdmissionController:
    """FIFO admission by device footprint and batch slots (host logic only).

    A pending request is admitted when the arena can place its footprint and a slot is free.
    A request that can never fit is rejected at submission. Requests behind a blocked head may
    pass it, at most ``max_bypass`` times in a row; then the queue waits for the head
    (no starvation).
    """

    def __init__(self, arena: DeviceArena, max_slots: int = 256, max_bypass: int = 16,
                 max_spins: int | None = None):
        if max_slots <= 0:
            raise ValueError("max_slots must be positive")
        self.arena = arena
        self.max_slots = int(max_slots)
        self.max_bypass = int(max_bypass)
        self.max_spins = max_spins
        self.pending: deque[Request] = deque()
        self.active: dict[int, tuple[Request, int]] = {}  # rid -> (request, arena offset)
        self._bypassed = 0

    def submit(self, req: Request) -> None:
        req.validate()
        if req.footprint() > self.arena.size:
            raise RequestRejected(f"request {req.rid} needs {req.footprint()} B of device "
                                  f"memory; the server holds {self.arena.size} B")
        if self.max_spins is not None and req.n_spins > self.max_spins:
            raise RequestRejected(f"request {req.rid} has {req.n_spins} spins; the server "
                                  f"takes at most {self.max_spins}")
        if req.rid in self.active or any(r.rid == req.rid for r in self.pending):
            raise RequestRejected(f"request id {req.rid} is already queued or running")
        self.pending.append(req)

    def cancel(self, rid: int) -> bool:
        """Drop a queued request; False if it is not queued (running or unknown)."""
        for r in self.pending:
            if r.rid == rid:
                self.pending.remove(r)
                return True
        return False

    def _try(self, req: Request) -> int | None:
        if len(self.active) >= self.max_slots:
            return None
        off = self.arena.alloc(req.footprint())
        if off is not None:
            self.active[req.rid] = (req, off)
        return off

    def admit(self) -> list[tuple[Request, int]]:
        out = []
        while self.pending and len(self.active) < self.max_slots:
            head = self.pending[0]
            off = self._try(head)
            if off is not None:
                self.pending.popleft()
                out.append((head, off))
                self._bypassed = 0
                continue
            if self._bypassed >= self.max_bypass:
                break
            for r in list(self.pending)[1:]:
                off = self._try(r)
                if off is not None:
                    self.pending.remove(r)
                    out.append((r, off))
                    self._bypassed += 1
                    break
            else:
                break
        return out

    def release(self, rid: int) -> Request:
        req, off = self.active.pop(rid)
        self.arena.release(off)
        return req


def poisson_requests(n: int, rate: float = 200.0, seed: int = 0, sweep_range=(100, 10_000),
                     sizes_2d=tuple(2**k for k in range(5, 13)), sizes_3d=(16, 32, 64, 128, 256),
                     p2d: float = 0.7) -> list[Request]:
    """A reproducible request stream: Poisson arrivals at ``rate`` per second; 2D with
    probability ``p2d`` (L uniform over ``sizes_2d``, int8 or msc with equal odds, msc only
    where L % 128 == 0, else int8), otherwise 3D int8 (L uniform over ``sizes_3d``, Metropolis
    or heat bath with equal odds); T uniform in [0.8, 1.2] T_c of the dimension; sweeps
    log-uniform in ``sweep_range``; a 63-bit seed per request."""
    rng = np.random.default_rng(seed)
    t = 0.0
    lo, hi = math.log(sweep_range[0]), math.log(sweep_range[1])
    out = []
    for i in range(n):
        t += rng.exponential(1.0 / rate)
        if rng.random() < p2d:
            dim, L = 2, int(rng.choice(sizes_2d))
            variant = "msc" if (rng.random() < 0.5 and L % 128 == 0) else "int8"
            alg, tc = "metropolis", TC_2D
        else:
            dim, L = 3, int(rng.choice(sizes_3d))
            variant = "int8"
            alg = "metropolis" if rng.random() < 0.5 else "heatbath"
            tc = 1.0 / BETA_C_3D
        T = float(rng.uniform(0.8, 1.2)) * tc
        sweeps = int(round(math.exp(rng.uniform(lo, hi))))
        out.append(Request(rid=i, dim=dim, L=L, variant=variant, algorithm=alg, T=T,
                           sweeps=sweeps, seed=int(rng.integers(0, 2**63)), arrival=t))
    return out


def make_simulation(req: Request, **kw):
    """The hip simulation of a request: hot start from numpy ``default_rng(seed)`` draws."""
    from .backends.hip_backend import HipInt8, HipMultiSpin

    cls = HipMultiSpin if req.variant == "msc" else HipInt8
    return cls(req.shape, 1.0 / req.T, seed=req.seed, algorithm=req.algorithm, init="hot", **kw)


def run_alone(req: Request, lattice: bool = False):
    """The request on its own through the plain single-lattice path: (M, E[, lattice])."""
    sim = make_simulation(req)
    sim.sweep(req.sweeps)
    M, E = sim.observables_exact()
    return (M, E, sim.get_lattice()) if lattice else (M, E)


class _Region:
    """Bump allocator over one arena region: hands out 256-byte aligned device views."""

    def __init__(self, base: int, size: int):
        self.base, self.size, self.used = int(base), int(size), 0

    def __call__(self, nbytes: int):
        from ._lib import DeviceView

        n = _align(nbytes)
        if self.used + n > self.size:
            raise MemoryError("simulation allocates more than its admitted footprint")
        v = DeviceView(self.base + self.used, nbytes)
        self.used += n
        return v


@dataclass
class _Active:
    req: Request
    sim: object
    remaining: int
    admitted: float


@dataclass
class ServeReport:
    results: list[Result] = field(default_factory=list)
    rejected: list[int] = field(default_factory=list)
    wall: float = 0.0
    steps: int = 0
    max_active: int = 0
    admitted_during_steps: int = 0

    def latencies(self) -> np.ndarray:
        return np.array([r.latency for r in self.results])

    def flips(self) -> int:
        return sum(r.sweeps * r.n_spins for r in self.results)

    def summary(self) -> dict:
        lat = self.latencies()
        q = np.array([r.admitted - r.arrival for r in self.results])
        return {"requests": len(self.results), "rejected": len(self.rejected),
                "p50_latency_ms": round(1e3 * float(np.percentile(lat, 50)), 3) if lat.size else None,
                "p99_latency_ms": round(1e3 * float(np.percentile(lat, 99)), 3) if lat.size else None,
                "p50_queue_ms": round(1e3 * float(np.percentile(q, 50)), 3) if q.size else None,
                "flips_per_ns": round(self.flips() / self.wall / 1e9, 2) if self.wall else None,
                "wall_s": round(self.wall, 3), "steps": self.steps,
                "max_active": self.max_active,
                "admitted_during_steps": self.admitted_during_steps}


class Server:
    """Continuous-batching executor over the hip backend.

    ``budget_bytes``: device memory for request state (one allocation, made here);
    ``max_slots``: lattices per batched step; ``step_sweeps``: sweeps per step (shorter steps
    let arrivals join sooner); ``max_spins``: largest request accepted (sizes the upload
    staging buffer); ``msc_mode``: msc tile kind of the batched kernel (0 compaction, 1
    lockstep: the default, since served msc lattices are small, at most ``max_spins``, and their
    rows of few words leave compaction tiles short of full waves).
    """

    def __init__(self, budget_bytes: int = 4 << 30, max_slots: int = 256, step_sweeps: int = 25,
                 max_spins: int = 1 << 24, msc_mode: int = 1, max_bypass: int = 16):
        from . import _lib
        from .batch import ReservedBatch

        self._lib = _lib
        self.arena = DeviceArena(budget_bytes)
        self.ctrl = AdmissionController(self.arena, max_slots, max_bypass, max_spins)
        self.step_sweeps = int(step_sweeps)
        self.setup = _lib.Stream()
        self.work = _lib.Stream()
        self._mem = _lib.DeviceBuffer(self.arena.size)
        self.staging = _lib.DeviceBuffer(int(max_spins))
        # a tile covers at least 2 KiB of colour-array state (int8: 256 threads x 4 sites x 2
        # colours), and each lattice adds at most one partial tile
        max_tiles = self.arena.size // 2048 + int(max_slots)
        self.batch = ReservedBatch(max_slots, max_tiles, self.work, msc_mode=msc_mode)

    def close(self) -> None:
        """Release the batch, the streams and the device memory now (explicitly, not at
        interpreter shutdown)."""
        self.batch.close()
        self.setup.sync()
        self._mem.free()
        self.staging.free()
        self.setup.close()
        self.work.close()

    def run(self, requests, realtime: bool = True, timeout: float = 600.0,
            keep_lattices: bool = False) -> ServeReport:
        """Serve ``requests``. ``realtime``: each request becomes visible at its arrival offset
        (seconds after the call); otherwise all are submitted at once. ``keep_lattices``: each
        result also carries the final full lattice."""
        L = self._lib
        rep = ServeReport()
        todo = deque(sorted(requests, key=lambda r: r.arrival))
        active: dict[int, _Active] = {}
        members: list[int] | None = None
        step_ev = L.Event()
        t0 = time.perf_counter()

        def clock() -> float:
            return time.perf_counter() - t0

        def intake() -> int:
            now = clock()
            while todo and (not realtime or todo[0].arrival <= now):
                r = todo.popleft()
                try:
                    self.ctrl.submit(r)
                except RequestRejected:
                    rep.rejected.append(r.rid)
            admitted = self.ctrl.admit()
            for r, off in admitted:
                sim = make_simulation(r, stream=self.setup.ptr,
                                      allocator=_Region(self._mem.ptr + off, r.footprint()),
                                      staging=self.staging)
                active[r.rid] = _Active(r, sim, r.sweeps, clock())
            return len(admitted)

        def retire(rid: int) -> None:
            a = active.pop(rid)
            M, E = a.sim.observables_exact()
            lat = a.sim.get_lattice() if keep_lattices else None
            r = a.req
            # latency counts from the request's arrival; in batch mode every request is
            # submitted at time 0
            arrival = r.arrival if realtime else 0.0
            rep.results.append(Result(r.rid, M, E, r.n_spins, r.sweeps, arrival, a.admitted,
                                      clock(), lat))
            self.ctrl.release(rid)

        while True:
            if clock() > timeout:
                raise TimeoutError(f"serving run exceeded {timeout} s")
            intake()
            for rid in [k for k, a in active.items() if a.remaining == 0]:
                retire(rid)
            if not active:
                if not todo and not self.ctrl.pending:
                    break
                if todo and not self.ctrl.pending:
                    time.sleep(max(0.0, min(todo[0].arrival - clock(), 1e-3)))
                continue
            ids = sorted(active)
            if ids != members:
                # each lattice is limited to its own remaining sweeps: it sits out the rest of
                # a step once done, so a step need not be cut to the shortest request
                self.batch.assign([active[i].sim for i in ids],
                                  [active[i].remaining for i in ids])
                members = ids
            step = min(self.step_sweeps, max(active[i].remaining for i in ids))
            self.batch.sweep(step)
            step_ev.record(self.work.ptr)
            rep.steps += 1
            rep.max_active = max(rep.max_active, len(ids))
            # while the step runs: take arrivals and prepare them on the setup stream
            while not step_ev.done():
                if realtime and todo:
                    rep.admitted_during_steps += intake()
                    if todo and todo[0].arrival > clock():
                        time.sleep(5e-5)
                else:
                    rep.admitted_during_steps += intake()
                    step_ev.sync()
                    break
            for i in ids:
                active[i].remaining -= min(step, active[i].remaining)
        self.setup.sync()
        rep.wall = clock()
        step_ev.close()
        return rep
