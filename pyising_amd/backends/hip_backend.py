"""The ``hip`` backend: gfx950 HIP kernels behind the C ABI of ``libising_hip.so``.

Drop-in for the reference's GPU variants (src/pyising/backends/cupy_backend.py:46-180):
``HipInt8`` ~ ``CupyInt8`` (int8 compressed checkerboard, 2D/3D, Metropolis and heat bath),
``HipMultiSpin`` ~ ``CupyMultiSpin`` (64 spins per uint64, 2D Metropolis, lazy comparison),
``HipMultiSpinEager`` ~ ``CupyMultiSpinEager`` (all 32 bits always drawn).

Differences by design: the hot start reproduces numpy ``default_rng(seed)`` draws exactly,
generated on the device from numpy's seeded PCG64 state (the reference's GPU backends use other
generators there), full-lattice <-> colour-array conversions run on the GPU, ``sweep(n)`` issues all 2n launches from C, and observables are
exact integer reductions on the device.
"""

from __future__ import annotations

import ctypes as C

import numpy as np

from .. import _lib
from ..lattice import pcg64_state, split_seed
from .base import IsingSimulation

_ONES_INT8 = 0x0101010101010101
_ALL_ONES = 0xFFFFFFFFFFFFFFFF

#: msc kernel modes of ising_msc_sweeps_ex / ising_msc_sweeps_sync
MSC_DEFAULT, MSC_COMPACT, MSC_LOCKSTEP, MSC_EAGER, MSC_BY_SIZE, MSC_FUSED = -1, 0, 1, 2, 3, 4
#: msc: resident sweeps (one cooperative launch per call, bands kept in LDS), a mode of the
#: Python class only (the others are the C library's)
MSC_RESIDENT = 5
#: int8 sweep forms: by size (resident sweeps for the 2D lattices they apply to), one launch
#: per half sweep, or resident sweeps (one cooperative launch per call, bands kept in LDS)
INT8_DEFAULT, INT8_HALVES, INT8_RESIDENT = -1, 0, 1
#: ISING_OBS_SCRATCH_BYTES of ising_hip.h: partial sums of the observable reductions
OBS_SCRATCH_BYTES = 16384
#: the observables result: (M, E) as int64 and a status word read back with them
OBS_OUT_BYTES = 24


def _full_lattice(lat: np.ndarray, shape) -> np.ndarray:
    a = np.asarray(lat)
    if a.size != int(np.prod(shape)):
        raise ValueError(f"lattice of {a.size} sites does not match shape {shape}")
    return np.ascontiguousarray(a, dtype=np.int8).reshape(shape)


class _HipBase(IsingSimulation):
    """Keyword-only options beyond the reference constructor (for servers that manage memory
    and streams themselves; the defaults give the plain drop-in behaviour):

    ``stream``: a created HIP stream (pointer) for every launch and copy of this simulation
    (default: the null stream). ``allocator``: callable ``nbytes -> device buffer`` for the
    resident state (colour arrays, reduction scratch; default: one hipMalloc each).
    ``staging``: a device buffer reused for host uploads of full lattices when it is large
    enough (default: a temporary allocation per upload).
    """

    backend = "hip"
    gpu = True

    def __init__(self, shape, beta: float, seed: int = 0, algorithm: str = "metropolis",
                 init="hot", *, stream=None, allocator=None, staging=None) -> None:
        self.stream_ptr = stream
        self._allocator = allocator
        self._staging = staging
        super().__init__(shape, beta, seed=seed, algorithm=algorithm, init=init)

    def _alloc(self, nbytes: int):
        return self._allocator(nbytes) if self._allocator else _lib.DeviceBuffer(nbytes)

    def _full_buffer(self, nbytes: int):
        """A device buffer for one full lattice: (staging or a new allocation, temporary?)."""
        if self._staging is not None and self._staging.nbytes >= nbytes:
            return self._staging, False
        return _lib.DeviceBuffer(nbytes), True

    def _upload_full(self, full: np.ndarray):
        """Full host lattice -> device; returns (buffer, temporary?)."""
        buf, tmp = self._full_buffer(full.nbytes)
        buf.upload(full, self.stream_ptr)
        return buf, tmp

    def _hot_start_full(self):
        """The hot start's {0, 1} draws (numpy ``default_rng(seed).integers(0, 2, shape, int8)``,
        src/pyising/lattice.py:93-96) generated on the device into a full-lattice buffer:
        the host only seeds numpy's PCG64; returns (buffer, temporary?)."""
        n = self.n_spins
        buf, tmp = self._full_buffer(n)
        st, inc = pcg64_state(self.seed)
        m64 = (1 << 64) - 1
        _lib.check(self._L.ising_hot_start(buf.ptr, n, st >> 64, st & m64, inc >> 64, inc & m64,
                                           self.stream_ptr), "hot_start")
        return buf, tmp

    def _hip_setup(self) -> None:
        self._L = _lib.lib()
        self.k0, self.k1 = split_seed(self.seed)
        self._obs_scratch = self._alloc(OBS_SCRATCH_BYTES)
        self._obs_out = self._alloc(OBS_OUT_BYTES)

    @staticmethod
    def state_bytes(shape, variant: str) -> list[int]:
        """Sizes of the device allocations a simulation of ``shape`` makes through its
        allocator: the reduction scratch, the (M, E) result and two colour arrays."""
        n = int(np.prod(shape))
        colour = n // 16 if variant == "msc" else n // 2
        return [OBS_SCRATCH_BYTES, OBS_OUT_BYTES, colour, colour]

    #: the batch (BatchScan or ReservedBatch) this lattice belongs to, while it does
    _batch = None

    def synchronize(self) -> None:
        _lib.check(self._L.ising_stream_sync(self.stream_ptr), "stream_sync")

    def _after_batch(self) -> None:
        """Order this simulation's next device work after the last batched sweep of the batch
        it belongs to: the batch runs on a stream of its own, which the simulation's stream
        (the null stream by default) does not otherwise wait for."""
        b = self._batch
        if b is not None:
            b.order_after(self.stream_ptr)

    def _own_sweep_allowed(self) -> None:
        if self._batch is not None:
            raise RuntimeError("this lattice belongs to a batch: sweep it through the batch "
                               "(the batch holds its sweep counter) or close the batch first")

    def close(self) -> None:
        """Wait for this simulation's work and release its device state now (colour arrays and
        reduction buffers; through the allocator's buffers when one was given). The object is
        unusable afterwards."""
        bufs = [getattr(self, "_obs_scratch", None), getattr(self, "_obs_out", None)]
        bufs += list(getattr(self, "c", None) or []) + list(getattr(self, "w", None) or [])
        if any(b is not None and getattr(b, "ptr", None) for b in bufs):
            self.synchronize()
        for b in bufs:
            if b is not None:
                b.free()

    def measure_series(self, n_meas: int, every: int = 1) -> np.ndarray:
        """``n_meas`` exact ``(M, E)`` pairs, one after every ``every`` sweeps, as an
        ``(n_meas, 2)`` int64 array: the same numbers as calling ``sweep(every)`` then
        ``observables_exact()`` in a loop, but each pair goes to its own device slot and all
        are read back once at the end, so the host never waits between measurements."""
        n_meas, every = int(n_meas), int(every)
        out = np.zeros((max(n_meas, 0), 2), dtype=np.int64)
        if n_meas <= 0:
            return out
        buf = _lib.DeviceBuffer(16 * n_meas)
        for i in range(n_meas):
            self.sweep(every)
            self._obs_into(buf.ptr + 16 * i)
        buf.download(out, self.stream_ptr)
        self.synchronize()
        buf.free()
        return out

    def _read_obs(self) -> tuple[int, int]:
        out = np.zeros(2, dtype=np.int64)
        self._obs_out.download(out, self.stream_ptr)
        return int(out[0]), int(out[1])

    def _read_obs_with_status(self, status_ptr) -> tuple[int, int, int]:
        """(M, E), already computed into the result buffer, and the 32-bit status word at
        ``status_ptr`` (a device copy into the buffer's third slot), in one read-back: one host
        wait where reading the status on its own would add a second."""
        _lib.check(self._L.ising_memcpy_d2d(self._obs_out.ptr + 16, status_ptr, 4,
                                            self.stream_ptr), "memcpy_d2d")
        out = np.zeros(3, dtype=np.int64)
        self._obs_out.download(out, self.stream_ptr)
        return int(out[0]), int(out[1]), int(out[2]) & 0xFFFFFFFF


class HipInt8(_HipBase):
    """int8 compressed checkerboard; V sites per lane, one Philox call per 4 sites."""

    variant = "int8"
    #: sweep form (INT8_DEFAULT, INT8_HALVES, INT8_RESIDENT) and, for resident sweeps, the
    #: number of bands (0: the library's choice); set before the first sweep
    mode = INT8_DEFAULT
    resident_bands = 0
    #: by size: resident sweeps from this many sweeps per call, for lattices up to this size
    #: (resident sweeps need rows of at most 64 KiB of LDS per band: L <= 2048 here)
    RESIDENT_MIN_SWEEPS = 2
    RESIDENT_MAX_SPINS = 2048 * 2048

    def _setup(self) -> None:
        self._hip_setup()
        self.ly = self.shape[-2]
        self.lz = self.shape[0] if self.ndim == 3 else 1
        nbytes = self.rows * self.lh
        self.c = [self._alloc(nbytes), self._alloc(nbytes)]
        self._table_arr = np.ascontiguousarray(self.table, dtype=np.uint32)
        self._rs = None  # workspace of the resident sweeps, made at their first use
        self._rs_bytes = None
        self._rs_epoch = C.c_uint32(0)
        self._rs_unchecked = False

    def resident_applies(self) -> bool:
        """Whether resident sweeps can run this lattice on this device (2D, lh % 16 == 0,
        bands and LDS within limits, a known workgroup placement)."""
        if self.ndim != 2:
            return False
        if self._rs_bytes is None:
            n = C.c_size_t()
            _lib.check(self._L.ising_int8_resident_bytes(self.rows, self.lh,
                                                         int(self.resident_bands), C.byref(n)),
                       "int8_resident_bytes")
            self._rs_bytes = n.value
        return self._rs_bytes > 0

    def _resident(self, n: int) -> bool:
        if self.mode == INT8_HALVES:
            return False
        if self.mode == INT8_RESIDENT:
            if not self.resident_applies():
                raise ValueError("resident int8 sweeps do not apply to this lattice here")
            return True
        return (n >= self.RESIDENT_MIN_SWEEPS and self.n_spins <= self.RESIDENT_MAX_SPINS
                and self.resident_applies())

    def check_resident(self) -> None:
        """Raise if a resident launch since the last check failed (a bounded wait for a
        neighbour band gave up); the lattice is then invalid and the workspace is zeroed again.
        Synchronous."""
        if not self._rs_unchecked or self._rs is None:
            return
        failed = C.c_int()
        _lib.check(self._L.ising_int8_resident_status(self._rs.ptr, C.byref(failed),
                                                      self.stream_ptr), "resident_status")
        self._rs_unchecked = False
        if failed.value:
            self._resident_failed(failed.value)

    def _resident_failed(self, status: int) -> None:
        """Zero the resident workspace again and raise (a resident launch reported `status`)."""
        _lib.check(self._L.ising_fill_u64(self._rs.ptr, self._rs.nbytes // 8, 0,
                                          self.stream_ptr), "fill")
        self._rs_epoch.value = 0
        raise _lib.HipError(f"resident int8 sweep failed (status {status}): the "
                            "lattice is invalid; sweep with mode=INT8_HALVES instead")

    def synchronize(self) -> None:
        super().synchronize()
        self.check_resident()

    def close(self) -> None:
        try:
            super().close()  # waits, and reports a resident failure not yet checked
        finally:
            if getattr(self, "_rs", None) is not None:
                self._rs.free()
                self._rs = None

    def _upload_table(self) -> None:
        self._table_arr = np.ascontiguousarray(self.table, dtype=np.uint32)

    def kernel(self) -> dict:
        """The per-half-sweep kernel's choice for this lattice: {"vec": sites per lane,
        "layout": 0 | 1 | 2, "name"} (resident sweeps, where they run, do not use it)."""
        vec, lay = C.c_int(), C.c_int()
        _lib.check(self._L.ising_int8_kernel(self.rows, self.lh, self.ly, self.lz, self.ndim,
                                             C.byref(vec), C.byref(lay)), "int8_kernel")
        names = {0: "per-lane rows", 1: "row waves", 2: "segment waves"}
        return {"vec": vec.value, "layout": lay.value, "name": names[lay.value]}

    def _split(self, buf, tmp: bool) -> None:
        self._after_batch()
        _lib.check(self._L.ising_split_int8(buf.ptr, self.c[0].ptr, self.c[1].ptr, self.rows,
                                            self.ly, self.shape[-1], self.stream_ptr),
                   "split_int8")
        self.synchronize()
        if tmp:
            buf.free()

    def set_lattice(self, lat: np.ndarray) -> None:
        self._split(*self._upload_full(_full_lattice(lat, self.shape)))

    def _fill_hot(self) -> None:
        self._split(*self._hot_start_full())

    def _fill_cold(self) -> None:
        n = self.rows * self.lh // 8
        for buf in self.c:
            _lib.check(self._L.ising_fill_u64(buf.ptr, n, _ONES_INT8, self.stream_ptr), "fill")

    def get_lattice(self) -> np.ndarray:
        self._after_batch()
        self.check_resident()
        full = np.empty(self.shape, dtype=np.int8)
        buf, tmp = self._full_buffer(full.nbytes)
        _lib.check(self._L.ising_merge_int8(self.c[0].ptr, self.c[1].ptr, buf.ptr, self.rows,
                                            self.ly, self.shape[-1], self.stream_ptr),
                   "merge_int8")
        buf.download(full, self.stream_ptr)
        if tmp:
            buf.free()
        return full

    def _half_sweep(self, color: int) -> None:
        self._own_sweep_allowed()
        own, other = self.c[color], self.c[1 - color]
        _lib.check(self._L.ising_int8_half(
            own.ptr, other.ptr, self.rows, self.lh, self.ly, self.lz, self.ndim,
            int(self.algorithm == "heatbath"), self._table_arr.ctypes.data,
            self.t & 0xFFFFFFFF, color, self.k0, self.k1, self.stream_ptr), "int8_half")

    def sweep(self, n: int = 1) -> None:
        n = int(n)
        if n <= 0:
            return
        self._own_sweep_allowed()
        if self._resident(n):
            if self._rs is None:
                self._rs = _lib.DeviceBuffer(self._rs_bytes)
                _lib.check(self._L.ising_fill_u64(self._rs.ptr, self._rs_bytes // 8, 0,
                                                  self.stream_ptr), "fill")
                self._rs_epoch.value = 0
            _lib.check(self._L.ising_int8_sweeps_resident(
                self.c[0].ptr, self.c[1].ptr, self.rows, self.lh,
                int(self.algorithm == "heatbath"), self._table_arr.ctypes.data,
                self.t & 0xFFFFFFFF, n, self.k0, self.k1, int(self.resident_bands), self._rs.ptr,
                C.byref(self._rs_epoch), self.stream_ptr), "int8_sweeps_resident")
            self._rs_unchecked = True
        else:
            _lib.check(self._L.ising_int8_sweeps(
                self.c[0].ptr, self.c[1].ptr, self.rows, self.lh, self.ly, self.lz, self.ndim,
                int(self.algorithm == "heatbath"), self._table_arr.ctypes.data,
                self.t & 0xFFFFFFFF, n, self.k0, self.k1, self.stream_ptr), "int8_sweeps")
        self.t += n

    def _obs_into(self, out_ptr) -> None:
        self._after_batch()
        _lib.check(self._L.ising_int8_obs(
            self.c[0].ptr, self.c[1].ptr, self.rows, self.lh, self.ly, self.lz, self.ndim,
            self._obs_scratch.ptr, out_ptr, self.stream_ptr), "int8_obs")

    def observables_exact(self) -> tuple[int, int]:
        self._obs_into(self._obs_out.ptr)
        if self._rs_unchecked and self._rs is not None:
            # the resident launches' status comes back with (M, E)
            m, e, status = self._read_obs_with_status(self._rs.ptr)
            self._rs_unchecked = False
            if status:
                self._resident_failed(status)
            return m, e
        return self._read_obs()


class HipMultiSpin(_HipBase):
    """Multi-spin coding: 64 spins per uint64, bit-sliced neighbour count, lazy comparison."""

    variant = "msc"
    dims = (2,)
    algorithms = ("metropolis",)
    layout_multiple = 128
    stream = "msc"
    bytes_per_update = 3.0 / 8.0
    #: kernel mode and its parameter (compaction: 10 * cap + R, see ising_msc_config); the
    #: process default picks the lockstep tile up to 2^18 words per colour (4096 x 8192) and
    #: queue compaction above
    mode = MSC_DEFAULT
    tile_k = 42
    #: resident sweeps: the number of bands (0: the library's choice); by size (MSC_DEFAULT or
    #: MSC_BY_SIZE) they run calls of at least RESIDENT_MIN_SWEEPS sweeps on lattices up to
    #: RESIDENT_MAX_SPINS spins
    resident_bands = 0
    RESIDENT_MIN_SWEEPS = 2
    #: measured, flips/ns best per-half-sweep kernel / resident: 1024^2 113 / 176, 2048^2
    #: 429 / 672, 4096^2 1514 / 2036, 8192^2 3722 / 3683
    RESIDENT_MAX_SPINS = 4096 * 4096

    def _setup(self) -> None:
        self._hip_setup()
        self.nw = self.lh // 64
        self.n_words = self.rows * self.nw
        self.w = [self._alloc(8 * self.n_words), self._alloc(8 * self.n_words)]
        self.stats_buf = None
        self._sync = None  # workspace of the fused sweeps, made at the first sweep
        self._epoch = C.c_uint32(0)
        self._sync_unchecked = False
        self._rs = None  # workspace of the resident sweeps, made at their first use
        self._rs_bytes = None
        self._rs_epoch = C.c_uint32(0)
        self._rs_unchecked = False

    def resident_applies(self) -> bool:
        """Whether resident sweeps can run this lattice on this device (an even word count per
        row, bands and LDS within limits, a known workgroup placement)."""
        if self._rs_bytes is None:
            n = C.c_size_t()
            _lib.check(self._L.ising_msc_resident_bytes(self.rows, self.nw,
                                                        int(self.resident_bands), C.byref(n)),
                       "msc_resident_bytes")
            self._rs_bytes = n.value
        return self._rs_bytes > 0

    def _resident(self, n: int) -> bool:
        if self.stats_buf is not None:  # the Philox statistics come from the other kernels
            return False
        if self.mode == MSC_RESIDENT:
            if not self.resident_applies():
                raise ValueError("resident msc sweeps do not apply to this lattice here")
            return True
        return (self.mode in (MSC_DEFAULT, MSC_BY_SIZE) and n >= self.RESIDENT_MIN_SWEEPS
                and self.n_spins <= self.RESIDENT_MAX_SPINS and self.resident_applies())

    def check_resident(self) -> None:
        """Raise if a resident launch since the last check failed (a bounded wait for a
        neighbour band gave up); the lattice is then invalid and the workspace is zeroed again.
        Synchronous."""
        if not self._rs_unchecked or self._rs is None:
            return
        failed = C.c_int()
        _lib.check(self._L.ising_msc_resident_status(self._rs.ptr, C.byref(failed),
                                                     self.stream_ptr), "resident_status")
        self._rs_unchecked = False
        if failed.value:
            self._resident_failed(failed.value)

    def _resident_failed(self, status: int) -> None:
        """Zero the resident workspace again and raise (a resident launch reported `status`)."""
        _lib.check(self._L.ising_fill_u64(self._rs.ptr, self._rs.nbytes // 8, 0,
                                          self.stream_ptr), "fill")
        self._rs_epoch.value = 0
        raise _lib.HipError(f"resident msc sweep failed (status {status}): the "
                            "lattice is invalid; sweep with mode=MSC_COMPACT instead")

    def _sync_buffer(self):
        """The fused sweeps' workspace (per-tile completion counters), zeroed once; its counter
        value travels with it in ``_epoch``."""
        if self._sync is None:
            n = C.c_size_t()
            _lib.check(self._L.ising_msc_sync_bytes(self.rows, self.nw, C.byref(n)), "sync_bytes")
            self._sync = _lib.DeviceBuffer(n.value)
            _lib.check(self._L.ising_fill_u64(self._sync.ptr, n.value // 8, 0, self.stream_ptr),
                       "fill")
            self._epoch.value = 0
        return self._sync

    def check_sync(self) -> None:
        """Raise if a fused launch since the last check failed (a bounded dependency wait
        gave up, or a tile's counter is not where every counter should be); the lattice is then
        not a valid continuation and the workspace is zeroed again. Synchronous."""
        if not self._sync_unchecked or self._sync is None:
            return
        failed = C.c_int()
        _lib.check(self._L.ising_msc_sync_status(self._sync.ptr, self.rows, self.nw,
                                                 self._epoch.value, C.byref(failed),
                                                 self.stream_ptr), "sync_status")
        self._sync_unchecked = False
        if failed.value:
            n = self._sync.nbytes
            _lib.check(self._L.ising_fill_u64(self._sync.ptr, n // 8, 0, self.stream_ptr), "fill")
            self._epoch.value = 0
            raise _lib.HipError(f"fused msc sweep failed (status {failed.value}): the lattice "
                                "is invalid; sweep with mode=MSC_COMPACT instead")

    def synchronize(self) -> None:
        super().synchronize()
        self.check_sync()
        self.check_resident()

    def close(self) -> None:
        try:
            super().close()  # waits, and reports a fused or resident failure not yet checked
        finally:
            if self._sync is not None:
                self._sync.free()
            if getattr(self, "_rs", None) is not None:
                self._rs.free()
                self._rs = None

    def _split(self, buf, tmp: bool) -> None:
        self._after_batch()
        _lib.check(self._L.ising_split_pack(buf.ptr, self.w[0].ptr, self.w[1].ptr, self.rows,
                                            self.shape[-1], self.stream_ptr), "split_pack")
        self.synchronize()
        if tmp:
            buf.free()

    def set_lattice(self, lat: np.ndarray) -> None:
        self._split(*self._upload_full(_full_lattice(lat, self.shape)))

    def _fill_hot(self) -> None:
        self._split(*self._hot_start_full())

    def _fill_cold(self) -> None:
        for buf in self.w:
            _lib.check(self._L.ising_fill_u64(buf.ptr, self.n_words, _ALL_ONES, self.stream_ptr),
                       "fill")

    def get_lattice(self) -> np.ndarray:
        self._after_batch()
        full = np.empty(self.shape, dtype=np.int8)
        buf, tmp = self._full_buffer(full.nbytes)
        _lib.check(self._L.ising_unpack_merge(self.w[0].ptr, self.w[1].ptr, buf.ptr, self.rows,
                                              self.shape[-1], self.stream_ptr), "unpack_merge")
        buf.download(full, self.stream_ptr)
        if tmp:
            buf.free()
        self.check_sync()
        self.check_resident()
        return full

    def kernel(self) -> dict:
        """The kernel a half sweep of this lattice runs: {"mode": 0 | 1 | 2, "name", "k"}."""
        mode, k = C.c_int(), C.c_int()
        _lib.check(self._L.ising_msc_kernel(self.rows, self.nw, self.mode, self.tile_k,
                                            C.byref(mode), C.byref(k)), "msc_kernel")
        names = {MSC_COMPACT: "queue compaction", MSC_LOCKSTEP: "lockstep", MSC_EAGER: "eager",
                 MSC_FUSED: "fused queue compaction"}
        return {"mode": mode.value, "name": names[mode.value], "k": k.value}

    def get_words(self) -> tuple[np.ndarray, np.ndarray]:
        """The two packed colour arrays, (rows, Lx/128) uint64 each."""
        self._after_batch()
        out = []
        for buf in self.w:
            a = np.empty((self.rows, self.nw), dtype=np.uint64)
            buf.download(a, self.stream_ptr)
            out.append(a)
        return out[0], out[1]

    def _stats_ptr(self):
        return None if self.stats_buf is None else self.stats_buf.ptr

    def _half_sweep(self, color: int) -> None:
        self._run(1, color_only=color)

    def _run(self, n: int, color_only: int | None = None) -> None:
        self._own_sweep_allowed()
        thr4, thr8 = int(self.table[3]), int(self.table[4])
        if color_only is not None:
            own, other = self.w[color_only], self.w[1 - color_only]
            _lib.check(self._L.ising_msc_half_ex(
                own.ptr, other.ptr, self.rows, self.nw, self.t & 0xFFFFFFFF, color_only,
                self.k0, self.k1, thr4, thr8, self.mode, self.tile_k, self._stats_ptr(),
                self.stream_ptr), "msc_half")
            return
        if self._resident(n):
            if self._rs is None:
                self._rs = _lib.DeviceBuffer(self._rs_bytes)
                _lib.check(self._L.ising_fill_u64(self._rs.ptr, self._rs_bytes // 8, 0,
                                                  self.stream_ptr), "fill")
                self._rs_epoch.value = 0
            _lib.check(self._L.ising_msc_sweeps_resident(
                self.w[0].ptr, self.w[1].ptr, self.rows, self.nw, self.t & 0xFFFFFFFF, n,
                self.k0, self.k1, thr4, thr8, int(self.resident_bands), self._rs.ptr,
                C.byref(self._rs_epoch), self.stream_ptr), "msc_sweeps_resident")
            self._rs_unchecked = True
            return
        sync = self._sync_buffer()
        _lib.check(self._L.ising_msc_sweeps_sync(
            self.w[0].ptr, self.w[1].ptr, self.rows, self.nw, self.t & 0xFFFFFFFF, n,
            self.k0, self.k1, thr4, thr8, self.mode, self.tile_k, sync.ptr,
            C.byref(self._epoch), self._stats_ptr(), self.stream_ptr), "msc_sweeps")
        self._sync_unchecked = True

    def sweep(self, n: int = 1) -> None:
        n = int(n)
        if n <= 0:
            return
        self._run(n)
        self.t += n

    def _obs_into(self, out_ptr) -> None:
        self._after_batch()
        _lib.check(self._L.ising_msc_obs(self.w[0].ptr, self.w[1].ptr, self.rows, self.nw,
                                         self._obs_scratch.ptr, out_ptr, self.stream_ptr),
                   "msc_obs")

    def observables_exact(self) -> tuple[int, int]:
        self._obs_into(self._obs_out.ptr)
        if self._rs_unchecked and self._rs is not None:
            # the resident launches' status comes back with (M, E)
            m, e, status = self._read_obs_with_status(self._rs.ptr)
            self._rs_unchecked = False
            self.check_sync()
            if status:
                self._resident_failed(status)
            return m, e
        out = self._read_obs()
        self.check_sync()
        return out

    def philox_stats(self, sweeps: int = 2) -> dict:
        """Philox work of ``sweeps`` sweeps (diagnostic; advances the simulation).

        ``calls_per_word``: calls the comparison needs (the reference's mean);
        ``slots_per_word``: lane-slots the SIMDs execute per word (64 x iterations per wave);
        ``efficiency`` = calls / slots.
        """
        self.stats_buf = _lib.DeviceBuffer(32)
        _lib.check(self._L.ising_fill_u64(self.stats_buf.ptr, 4, 0, self.stream_ptr), "fill")
        self.sweep(sweeps)
        st = np.zeros(4, dtype=np.uint64)
        self.stats_buf.download(st, self.stream_ptr)
        self.stats_buf = None
        calls, slots, words = (int(x) for x in st[:3])
        return {"calls_per_word": calls / max(words, 1), "slots_per_word": slots / max(words, 1),
                "efficiency": calls / max(slots, 1), "words": words}


class HipMultiSpinEager(HipMultiSpin):
    """Ablation: identical results, but all 16 Philox calls per word are always issued."""

    variant = "msc-eager32"
    mode = MSC_EAGER
