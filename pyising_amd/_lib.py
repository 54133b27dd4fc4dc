"""ctypes binding of ``libising_hip.so``, the C ABI over the gfx950 kernels.

The library is built in-tree by :mod:`pyising_amd.build` (``__graft_entry__.build()``). Every
entry point returns 0 or a negative status; :func:`check` turns a failure into an exception
carrying ``ising_last_error()``. There is no CPU fallback: if the library is missing or no GPU
is visible, :func:`lib` raises.
"""

from __future__ import annotations

import atexit
import ctypes as C
import sys
import threading
from pathlib import Path

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "libising_hip.so"

_P = C.c_void_p
_I = C.c_int
_LL = C.c_longlong
_U32 = C.c_uint32
_U64 = C.c_uint64
_SZ = C.c_size_t

#: argument types of every exported entry point (mirrors include/ising_hip.h)
SIGNATURES: dict[str, list] = {
    "ising_device_count": [_P],
    "ising_device_info": [_P, _I, _P, _P],
    "ising_malloc": [_P, _SZ],
    "ising_free": [_P],
    "ising_mem_info": [_P, _P],
    "ising_memcpy_h2d": [_P, _P, _SZ, _P],
    "ising_memcpy_d2h": [_P, _P, _SZ, _P],
    "ising_memcpy_d2d": [_P, _P, _SZ, _P],
    "ising_stream_create": [_P],
    "ising_stream_destroy": [_P],
    "ising_stream_sync": [_P],
    "ising_device_sync": [],
    "ising_event_create": [_P],
    "ising_event_destroy": [_P],
    "ising_event_record": [_P, _P],
    "ising_event_query": [_P],
    "ising_event_elapsed_ms": [_P, _P, _P],
    "ising_event_sync": [_P],
    "ising_stream_wait_event": [_P, _P],
    "ising_philox_kat": [_P, _P, _P, _I],
    "ising_philox_rate": [_P, _I, _I, _U32, _U32, _P],
    "ising_copy": [_P, _P, _SZ, _I, _P],
    "ising_fill_u64": [_P, _LL, _U64, _P],
    "ising_hot_start": [_P, _LL, _U64, _U64, _U64, _U64, _P],
    "ising_split_int8": [_P, _P, _P, _LL, _I, _I, _P],
    "ising_merge_int8": [_P, _P, _P, _LL, _I, _I, _P],
    "ising_split_pack": [_P, _P, _P, _LL, _I, _P],
    "ising_unpack_merge": [_P, _P, _P, _LL, _I, _P],
    "ising_int8_half": [_P, _P, _LL, _I, _I, _I, _I, _I, _P, _U32, _U32, _U32, _U32, _P],
    "ising_int8_sweeps": [_P, _P, _LL, _I, _I, _I, _I, _I, _P, _U32, _LL, _U32, _U32, _P],
    "ising_int8_resident_bytes": [_LL, _I, _I, _P],
    "ising_int8_sweeps_resident": [_P, _P, _LL, _I, _I, _P, _U32, _LL, _U32, _U32, _I, _P, _P, _P],
    "ising_int8_resident_status": [_P, _P, _P],
    "ising_int8_kernel": [_LL, _I, _I, _I, _I, _P, _P],
    "ising_msc_resident_bytes": [_LL, _I, _I, _P],
    "ising_msc_sweeps_resident": [_P, _P, _LL, _I, _U32, _LL, _U32, _U32, _U32, _U32, _I, _P, _P, _P],
    "ising_msc_resident_status": [_P, _P, _P],
    "ising_int8_obs": [_P, _P, _LL, _I, _I, _I, _I, _P, _P, _P],
    "ising_msc_config": [_I, _I],
    "ising_msc_kernel": [_LL, _I, _I, _I, _P, _P],
    "ising_msc_half": [_P, _P, _LL, _I, _U32, _U32, _U32, _U32, _U32, _U32, _P, _P],
    "ising_msc_sweeps": [_P, _P, _LL, _I, _U32, _LL, _U32, _U32, _U32, _U32, _P, _P],
    "ising_msc_half_ex": [_P, _P, _LL, _I, _U32, _U32, _U32, _U32, _U32, _U32, _I, _I, _P, _P],
    "ising_msc_sweeps_ex": [_P, _P, _LL, _I, _U32, _LL, _U32, _U32, _U32, _U32, _I, _I, _P, _P],
    "ising_msc_obs": [_P, _P, _LL, _I, _P, _P, _P],
    "ising_msc_obs_static": [_P, _P, _LL, _I, _P, _P],
    "ising_msc_sweeps_sync": [_P, _P, _LL, _I, _U32, _LL, _U32, _U32, _U32, _U32, _I, _I, _P,
                              _P, _P, _P],
    "ising_msc_sync_bytes": [_LL, _I, _P],
    "ising_msc_sync_status": [_P, _LL, _I, _U32, _P, _P],
    "ising_batch_create": [_P, _I, _I, _I, _I, _P],
    "ising_batch_sweeps": [_P, _LL, _P],
    "ising_batch_graph": [_P, _LL, _P],
    "ising_batch_replay": [_P, _I, _P],
    "ising_batch_stats": [_P, _P, _P, _P],
    "ising_batch_reset_stats": [_P, _P],
    "ising_batch_info": [_P, _P, _P, _P],
    "ising_batch_set_persistent": [_P, _I],
    "ising_batch_status": [_P, _P, _P],
    "ising_batch_set_moments": [_P, _I],
    "ising_batch_set_tiles_per_workgroup": [_P, _I],
    "ising_batch_destroy": [_P],
    "ising_batch_tiles": [_P, _I, _I, _P],
    "ising_batch_reserve": [_I, _LL, _I, _I, _I, _P],
    "ising_batch_assign": [_P, _P, _I, _P],
}


class HipError(RuntimeError):
    """A C-ABI entry point returned a failure status."""


_lib: C.CDLL | None = None
_lock = threading.Lock()
_finalizing = False


def _mark_finalizing() -> None:
    global _finalizing
    _finalizing = True


# Runs when the interpreter starts to shut down, before module objects are collected: from
# then on garbage-collected handles release nothing. A release at that point could run after
# the HIP runtime (or a profiler's tool library) has begun its own teardown, and a native fault
# there cannot be caught. Programs release what they hold with the explicit close()/free()
# calls; the process's exit returns the rest to the driver.
atexit.register(_mark_finalizing)


def releasing_allowed() -> bool:
    """False once interpreter shutdown has begun: ``__del__`` paths must then call nothing in
    the library."""
    return not (_finalizing or sys.is_finalizing())


def load_library() -> C.CDLL:
    """Load the shared library and declare its signatures (no GPU needed)."""
    global _lib
    with _lock:
        if _lib is None:
            if not LIB_PATH.exists():
                raise ImportError(f"{LIB_PATH} is not built; run __graft_entry__.build() or "
                                  "python -m pyising_amd.build")
            lib = C.CDLL(str(LIB_PATH))
            for name, args in SIGNATURES.items():
                fn = getattr(lib, name)
                fn.argtypes = args
                fn.restype = _I
            lib.ising_last_error.argtypes = []
            lib.ising_last_error.restype = C.c_char_p
            _lib = lib
    return _lib


def check(rc: int, what: str = "") -> int:
    if rc < 0:
        msg = load_library().ising_last_error().decode(errors="replace")
        raise HipError(f"{what or 'ising call'} failed ({rc}): {msg}")
    return rc


def device_count() -> int:
    """Number of visible GPUs (0 when the runtime reports none or fails)."""
    n = C.c_int(0)
    try:
        rc = load_library().ising_device_count(C.byref(n))
    except (ImportError, OSError):
        return 0
    return n.value if rc == 0 else 0


def lib() -> C.CDLL:
    """The library, with at least one GPU visible; raises otherwise (no CPU fallback)."""
    L = load_library()
    if device_count() < 1:
        raise HipError("no HIP device visible: the hip backend runs only on a GPU")
    return L


class DeviceBuffer:
    """Owned device allocation (freed on garbage collection)."""

    __slots__ = ("ptr", "nbytes", "_lib")

    def __init__(self, nbytes: int):
        self._lib = lib()
        p = C.c_void_p()
        check(self._lib.ising_malloc(C.byref(p), int(nbytes)), "ising_malloc")
        self.ptr = p.value
        self.nbytes = int(nbytes)

    def free(self) -> None:
        if getattr(self, "ptr", None):
            self._lib.ising_free(self.ptr)
            self.ptr = None

    def __del__(self):
        if not releasing_allowed():
            return
        try:
            self.free()
        except Exception:  # noqa: BLE001 - a failing release in a finalizer is dropped
            pass

    def upload(self, host, stream=None) -> None:
        import numpy as np

        a = np.ascontiguousarray(host)
        if a.nbytes > self.nbytes:
            raise ValueError("upload larger than the buffer")
        check(self._lib.ising_memcpy_h2d(self.ptr, a.ctypes.data, a.nbytes, stream), "h2d")

    def download(self, host, stream=None) -> None:
        if host.nbytes > self.nbytes or not host.flags["C_CONTIGUOUS"]:
            raise ValueError("download target must be contiguous and fit the buffer")
        check(self._lib.ising_memcpy_d2h(host.ctypes.data, self.ptr, host.nbytes, stream), "d2h")


class DeviceView(DeviceBuffer):
    """A range of device memory owned elsewhere (a sub-allocation); ``free`` releases nothing
    but the view, or hands the range back through ``release``."""

    __slots__ = ("_release",)

    def __init__(self, ptr: int, nbytes: int, release=None):  # noqa: D107 - no allocation
        self._lib = lib()
        self.ptr = int(ptr)
        self.nbytes = int(nbytes)
        self._release = release

    def free(self) -> None:
        if getattr(self, "ptr", None):
            self.ptr = None
            if self._release is not None:
                self._release()


class Event:
    """HIP event (timing on a given stream)."""

    def __init__(self):
        self._lib = lib()
        p = C.c_void_p()
        check(self._lib.ising_event_create(C.byref(p)), "event_create")
        self.ptr = p.value

    def record(self, stream=None) -> None:
        check(self._lib.ising_event_record(self.ptr, stream), "event_record")

    def elapsed_ms(self, end: "Event") -> float:
        ms = C.c_float()
        check(self._lib.ising_event_elapsed_ms(C.byref(ms), self.ptr, end.ptr), "event_elapsed")
        return float(ms.value)

    def done(self) -> bool:
        return check(self._lib.ising_event_query(self.ptr), "event_query") == 1

    def sync(self) -> None:
        check(self._lib.ising_event_sync(self.ptr), "event_sync")

    def close(self) -> None:
        if getattr(self, "ptr", None):
            self._lib.ising_event_destroy(self.ptr)
            self.ptr = None

    def __del__(self):
        if not releasing_allowed():
            return
        try:
            self.close()
        except Exception:  # noqa: BLE001 - a failing release in a finalizer is dropped
            pass


class Stream:
    """Non-blocking HIP stream at the default priority."""

    def __init__(self):
        self._lib = lib()
        p = C.c_void_p()
        check(self._lib.ising_stream_create(C.byref(p)), "stream_create")
        self.ptr = p.value

    def sync(self) -> None:
        check(self._lib.ising_stream_sync(self.ptr), "stream_sync")

    def wait(self, event: Event) -> None:
        """Later work on this stream waits for the last record of ``event``."""
        check(self._lib.ising_stream_wait_event(self.ptr, event.ptr), "stream_wait_event")

    def close(self) -> None:
        """Wait for the stream's work, then destroy it."""
        if getattr(self, "ptr", None):
            self._lib.ising_stream_sync(self.ptr)
            self._lib.ising_stream_destroy(self.ptr)
            self.ptr = None

    def __del__(self):
        if not releasing_allowed():
            return
        try:
            self.close()
        except Exception:  # noqa: BLE001 - a failing release in a finalizer is dropped
            pass


def device_sync() -> None:
    check(lib().ising_device_sync(), "device_sync")


def mem_info() -> tuple[int, int]:
    """(free, total) bytes of the current device."""
    free, total = C.c_size_t(), C.c_size_t()
    check(lib().ising_mem_info(C.byref(free), C.byref(total)), "mem_info")
    return free.value, total.value
