"""Hardware ceilings and profiler targets (the counterpart of the reference's bench/roofline.py).

* :func:`copy_bandwidth` -- streaming copy with 4 GiB buffers (read + write bytes counted),
  the memory roofline; far larger than the 256 MB Infinity Cache.
* :func:`philox_rate` -- Philox4x32-10 calls per ns over the whole GPU, the RNG roofline
  (method of bench/roofline.py:24-36,59-78 of the reference: 512 calls per thread).

Command line (short, deterministic runs to put under ``rocprofv3``)::

    python -m pyising_amd.probes copy --mib 1024 --reps 4     # calibration for FETCH/WRITE_SIZE
    python -m pyising_amd.probes msc --L 16384 --sweeps 10     # msc half sweeps at T_c
    python -m pyising_amd.probes int8 --L 8192 --sweeps 10     # int8 2D half sweeps at T_c
    python -m pyising_amd.probes batch --sweeps 50             # Binder-scan batch, 128 lattices
    python -m pyising_amd.probes serve --requests 300          # serving loop, back to back
"""

from __future__ import annotations

import argparse
import ctypes as C
import hashlib
import json
import sys
from pathlib import Path

import numpy as np

from . import _lib


#: sources the msc half-sweep kernel is compiled from (a PMC summary records their digest, so
#: a reader can tell whether it was taken on the kernel built now)
MSC_KERNEL_SOURCES = ("common.hpp", "msc_core.hpp", "msc_kernels.hip", "msc_fused.hip")


def kernel_source_digest(names=MSC_KERNEL_SOURCES) -> str:
    """First 16 hex digits of sha256 over the named csrc files (in order) and the compile
    command's flags."""
    from .build import CSRC, FLAGS

    h = hashlib.sha256()
    for n in names:
        h.update((CSRC / n).read_bytes())
    h.update(" ".join(FLAGS).encode())
    return h.hexdigest()[:16]


def device_info() -> dict:
    lib = _lib.lib()
    name = C.create_string_buffer(64)
    n_cu = C.c_int()
    total = C.c_longlong()
    _lib.check(lib.ising_device_info(name, 64, C.byref(n_cu), C.byref(total)))
    return {"arch": name.value.decode(), "cu": n_cu.value, "hbm_bytes": total.value}


def copy_bandwidth(nbytes: int = 4 << 30, repeats: int = 20, blocks: int = 1024) -> float:
    """Median copy bandwidth in GB/s (bytes read + bytes written)."""
    lib = _lib.lib()
    a = _lib.DeviceBuffer(nbytes)
    b = _lib.DeviceBuffer(nbytes)
    e0, e1 = _lib.Event(), _lib.Event()
    _lib.check(lib.ising_copy(b.ptr, a.ptr, nbytes, blocks, None))
    rates = []
    for _ in range(repeats):
        e0.record()
        _lib.check(lib.ising_copy(b.ptr, a.ptr, nbytes, blocks, None))
        e1.record()
        rates.append(2 * nbytes / (e0.elapsed_ms(e1) * 1e-3) / 1e9)
    for h in (a, b):
        h.free()
    for ev in (e0, e1):
        ev.close()
    return float(np.median(rates))


def philox_rate(calls: int = 512, repeats: int = 7, blocks_per_cu: int = 32) -> float:
    """Median Philox4x32-10 calls per ns over the whole GPU."""
    lib = _lib.lib()
    blocks = device_info()["cu"] * blocks_per_cu
    out = _lib.DeviceBuffer(blocks * 256 * 4)
    e0, e1 = _lib.Event(), _lib.Event()
    _lib.check(lib.ising_philox_rate(out.ptr, blocks, calls, 1, 2, None))
    rates = []
    for _ in range(repeats):
        e0.record()
        _lib.check(lib.ising_philox_rate(out.ptr, blocks, calls, 1, 2, None))
        e1.record()
        rates.append(blocks * 256 * calls / (e0.elapsed_ms(e1) * 1e-3) / 1e9)
    out.free()
    for ev in (e0, e1):
        ev.close()
    return float(np.median(rates))


def _run_copy(mib: int, reps: int) -> dict:
    lib = _lib.lib()
    n = mib << 20
    a = _lib.DeviceBuffer(n)
    b = _lib.DeviceBuffer(n)
    for _ in range(reps):
        _lib.check(lib.ising_copy(b.ptr, a.ptr, n, 1024, None))
    _lib.device_sync()
    a.free()
    b.free()
    return {"probe": "copy", "bytes_read_per_launch": n, "bytes_written_per_launch": n,
            "launches": reps}


def _run_msc(L: int, sweeps: int, mode: int, k: int) -> dict:
    from .backends.hip_backend import HipMultiSpin
    from .lattice import BETA_C_2D

    sim = HipMultiSpin((L, L), BETA_C_2D, seed=1)
    sim.mode, sim.tile_k = mode, k
    sim.sweep(sweeps)
    sim.close()
    return {"probe": "msc", "L": L, "sweeps": sweeps, "launches": 2 * sweeps,
            "algorithmic_bytes_per_launch": int(0.375 * L * L / 2)}


def _run_int8(L: int, sweeps: int, dim: int, algorithm: str, mode: int = -1) -> dict:
    from .backends.hip_backend import HipInt8
    from .lattice import BETA_C_2D, BETA_C_3D

    beta = BETA_C_2D if dim == 2 else BETA_C_3D
    sim = HipInt8((L,) * dim, beta, seed=1, algorithm=algorithm)
    sim.mode = mode  # -1 by size, 0 one launch per half sweep, 1 resident sweeps
    sim.sweep(sweeps)
    sim.synchronize()
    sim.close()
    return {"probe": "int8", "L": L, "dim": dim, "sweeps": sweeps, "launches": 2 * sweeps,
            "algorithmic_bytes_per_launch": int(3.0 * L**dim / 2)}


def _run_batch(sweeps: int, msc_tile: str, moments: bool, persistent: bool = False,
               k: int = 1) -> dict:
    """The Binder-scan batch (96 int8 lattices L = 64..2048, 32 msc lattices L = 1024, 4096,
    16 temperatures each), fixed tiles (msc_tile "mixed": lockstep for L = 1024, queue
    compaction for L = 4096), no graph; persistent: dependency-ordered launches; k tiles per
    workgroup."""
    from .batch import BatchScan, temperature_scan_lattices

    sims = temperature_scan_lattices()
    if msc_tile == "mixed":
        tiles = {(1024, 1024): "lockstep", (4096, 4096): "queue"}
    else:
        tiles = {(1024, 1024): msc_tile, (4096, 4096): msc_tile}
    scan = BatchScan(sims, msc_tiles=tiles, moments=moments, persistent=persistent,
                     tiles_per_workgroup=k)
    scan.sweep(sweeps)
    scan.synchronize()
    scan.close()
    for sim in sims:
        sim.close()
    return {"probe": "batch", "lattices": len(sims), "spins": int(sum(s.n_spins for s in sims)),
            "sweeps": sweeps, "msc_tile": msc_tile, "moments": moments,
            "persistent": persistent, "tiles_per_workgroup": k}


def _run_serve(requests: int, rate: float) -> dict:
    """The serving loop on the Poisson request stream, served back to back (not real time)."""
    from .serve import Server, poisson_requests

    reqs = poisson_requests(requests, rate=rate, seed=0)
    srv = Server(budget_bytes=8 << 30, max_slots=256, step_sweeps=25, max_spins=1 << 24)
    rep = srv.run(reqs, realtime=False, timeout=600)
    srv.close()
    return {"probe": "serve", **rep.summary()}


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawTextHelpFormatter)
    sub = ap.add_subparsers(dest="cmd", required=True)
    c = sub.add_parser("copy")
    c.add_argument("--mib", type=int, default=1024)
    c.add_argument("--reps", type=int, default=4)
    m = sub.add_parser("msc")
    m.add_argument("--L", type=int, default=16384)
    m.add_argument("--sweeps", type=int, default=10)
    m.add_argument("--mode", type=int, default=0)
    m.add_argument("--k", type=int, default=42)
    i = sub.add_parser("int8")
    i.add_argument("--L", type=int, default=8192)
    i.add_argument("--dim", type=int, default=2)
    i.add_argument("--sweeps", type=int, default=10)
    i.add_argument("--algorithm", default="metropolis")
    i.add_argument("--mode", type=int, default=-1, choices=(-1, 0, 1))
    ph = sub.add_parser("philox")
    ph.add_argument("--calls", type=int, default=512)
    ph.add_argument("--reps", type=int, default=3)
    sub.add_parser("ceilings")
    sub.add_parser("digest")
    b = sub.add_parser("batch")
    b.add_argument("--sweeps", type=int, default=50)
    b.add_argument("--msc-tile", default="lockstep", choices=("lockstep", "queue", "mixed"))
    b.add_argument("--no-moments", action="store_true")
    b.add_argument("--persistent", action="store_true")
    b.add_argument("--k", type=int, default=1, choices=(1, 2, 4, 8))
    sv = sub.add_parser("serve")
    sv.add_argument("--requests", type=int, default=300)
    sv.add_argument("--rate", type=float, default=200.0)
    a = ap.parse_args(argv)
    if a.cmd == "copy":
        out = _run_copy(a.mib, a.reps)
    elif a.cmd == "philox":
        rate = philox_rate(calls=a.calls, repeats=a.reps)
        out = {"probe": "philox", "calls_per_thread": a.calls, "philox_calls_per_ns": rate}
    elif a.cmd == "msc":
        out = _run_msc(a.L, a.sweeps, a.mode, a.k)
    elif a.cmd == "int8":
        out = _run_int8(a.L, a.sweeps, a.dim, a.algorithm, a.mode)
    elif a.cmd == "batch":
        out = _run_batch(a.sweeps, a.msc_tile, not a.no_moments, a.persistent, a.k)
    elif a.cmd == "serve":
        out = _run_serve(a.requests, a.rate)
    elif a.cmd == "digest":
        out = {"kernel_sources_sha16": kernel_source_digest()}
    else:
        out = {"copy_bandwidth_GBs": copy_bandwidth(), "philox_calls_per_ns": philox_rate(),
               **device_info()}
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
