"""Benchmark: 2D multi-spin-coded Metropolis sweeps of an L=16384 torus at T_c on one MI355X.

python bench.py [--gpus 1] [--steps K] [--warmup W]

One step = one full checkerboard sweep (two half-sweep launches) of the 16384^2 lattice
(beta = 1/T_c exact, seed 1, hot start = numpy default_rng(1) draws), resident in HBM. Prints
one JSON line: flips/ns (sweeps x spins / s / 1e9), the roofline of the dominant kernel (the
msc half sweep) under the 0.375 B/update traffic model against the 8 TB/s HBM peak, the
measured copy bandwidth and Philox ceiling, the batched-scan comparison, the serving loop on the
Poisson request stream, and the oracle timed on one host core as the CPU baseline.
"""

from __future__ import annotations

import argparse
import gc
import json
import os
import platform
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

L_HEADLINE = 16384
HBM_PEAK_GBS = 8000.0
BYTES_PER_UPDATE = 0.375
PMC_FILE = ROOT / "profiles" / "msc_L16384_pmc.json"


def cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def _time_like_reference(step, min_repeat_s: float, repeats: int = 5):
    """The reference's time_one (bench/run.py:121-142): one warm-up call, then ``repeats``
    timed repeats of whole calls, each repeat at least ``min_repeat_s`` long; returns
    (calls per repeat, median seconds per call, total seconds timed)."""
    step()  # warm-up (allocations, page faults)
    t0 = time.perf_counter()
    step()
    one = time.perf_counter() - t0
    per_repeat = max(1, int(np.ceil(min_repeat_s / max(one, 1e-9))))
    times = []
    for _ in range(repeats):
        t0 = time.perf_counter()
        for _ in range(per_repeat):
            step()
        times.append((time.perf_counter() - t0) / per_repeat)
    return per_repeat, float(np.median(times)), float(sum(times) * per_repeat)


def cpu_baseline(min_repeat_s: float = 0.5) -> dict:
    """The oracle's NumPy msc sweep on the same lattice (L=16384, T_c, seed 1, hot start) on
    one host core, timed like the reference's bench (warm-up sweep, then the median of 5
    repeats of whole sweeps, each repeat >= ``min_repeat_s``; one sweep takes ~2.5 s here)."""
    from oracle import ising_oracle as O
    from pyising_amd.lattice import BETA_C_2D

    sim = O.MscOracle((L_HEADLINE, L_HEADLINE), BETA_C_2D, seed=1)
    per, sec, total = _time_like_reference(lambda: sim.sweep(1), min_repeat_s)
    rate = L_HEADLINE**2 / sec / 1e9
    out = {"value": round(rate, 4), "unit": "flips/ns", "cores": 1, "kind": "port",
           "sample": f"oracle NumPy msc sweep, L={L_HEADLINE} T_c seed 1 hot start: warm-up "
                     f"sweep, then median of 5 repeats of {per} sweep(s) ({total:.1f} s timed) "
                     f"on one core of {cpu_model()} (os.cpu_count()={os.cpu_count()})"}
    try:
        out["openmp_c"] = cpu_baseline_openmp(min_repeat_s)
    except (OSError, RuntimeError) as exc:  # the C restatement was not built
        out["openmp_c"] = {"error": str(exc)[:200]}
    return out


def cpu_baseline_openmp(min_repeat_s: float = 0.5, threads: int = 16) -> dict:
    """The oracle's C restatement of the same msc sweep (same lattice and stream) with OpenMP
    over rows on ``threads`` host threads (the box's CPU share), timed like the NumPy one."""
    from oracle import c_oracle as CO
    from pyising_amd.lattice import BETA_C_2D

    threads = max(1, min(threads, os.cpu_count() or 1))
    CO.set_threads(threads)
    sim = CO.CMsc((L_HEADLINE, L_HEADLINE), BETA_C_2D, seed=1)
    per, sec, total = _time_like_reference(lambda: sim.sweep(1), min_repeat_s)
    rate = L_HEADLINE**2 / sec / 1e9
    return {"value": round(rate, 4), "unit": "flips/ns", "cores": threads, "kind": "port",
            "sample": f"oracle C msc sweep (OpenMP over rows), L={L_HEADLINE} T_c seed 1 hot "
                      f"start: warm-up sweep, then median of 5 repeats of {per} sweeps "
                      f"({total:.1f} s timed) on {threads} threads of {cpu_model()}"}


def batch_vs_plain(batch_sweeps: int = 200, plain_sweeps: int = 20, repeats: int = 3) -> dict:
    """Binder temperature scan: 96 int8 lattices (L = 64..2048) and 32 msc lattices
    (L = 1024, 4096), 16 temperatures in [2.20, 2.34] each, seeds base + index, hot starts.
    Aggregate flips/ns of the batched scan (all lattices per launch; tiles, launch form and graph
    length picked by the autotuner, its timing outside the measured region) versus the plain
    per-lattice loop (``sim.sweep(n)`` lattice after lattice). Each side runs its timed work
    once untimed first (same kernels, workspaces made, clocks up), then reports the median of
    ``repeats`` timed runs."""
    from pyising_amd import _lib
    from pyising_amd.batch import BatchScan, temperature_scan_lattices

    sims = temperature_scan_lattices()
    n_spins = sum(s.n_spins for s in sims)
    scan = BatchScan.tuned(sims)  # autotuner: msc tile per L, launch form, graph length
    per_replay = scan.graph_sweeps
    graph = per_replay > 0
    if not graph:
        per_replay = 50
    reps = max(1, batch_sweeps // per_replay)
    e0, e1 = _lib.Event(), _lib.Event()

    def batch_run() -> None:
        if graph:
            scan.replay(reps)
        else:
            scan.sweep(reps * per_replay)

    batch_run()
    scan.synchronize()
    rates = []
    for _ in range(repeats):
        e0.record(scan.stream.ptr)
        batch_run()
        e1.record(scan.stream.ptr)
        rates.append(reps * per_replay * n_spins / (e0.elapsed_ms(e1) * 1e-3) / 1e9)
    batch_rate = float(np.median(rates))
    mom = scan.moments()
    tuning = scan.tuning
    scan.close()

    def plain_run() -> None:
        for s in sims:
            s.sweep(plain_sweeps)
        _lib.device_sync()

    plain_run()
    times = []
    for _ in range(repeats):
        t0 = time.perf_counter()
        plain_run()
        times.append(time.perf_counter() - t0)
    plain_rate = plain_sweeps * n_spins / float(np.median(times)) / 1e9
    for h in list(sims) + [e0, e1]:
        h.close()
    return {"lattices": len(sims), "spins": int(n_spins), "graph": bool(graph),
            "batched_flips_per_ns": round(batch_rate, 1), "plain_flips_per_ns": round(plain_rate, 1),
            "speedup": round(batch_rate / plain_rate, 2),
            "timing": f"median of {repeats} timed runs each ({reps * per_replay} batch sweeps, "
                      f"{plain_sweeps} sweeps per lattice in the plain loop) after one untimed run",
            "tuning": {"msc_tiles": {"x".join(map(str, k)): v for k, v in tuning.msc_tiles.items()},
                       "int8_vecs": {"x".join(map(str, k)): v for k, v in tuning.int8_vecs.items()},
                       "persistent": tuning.persistent, "graph_sweeps": tuning.graph_sweeps,
                       "tiles_per_workgroup": tuning.tiles_per_workgroup},
            # a few hundred sweeps after a hot start: a check that the device moment sums are
            # read back, not an equilibrated Binder cumulant (validate.py measures those)
            "moments_L4096_T2.265": {"samples": mom[-9].samples, "binder": round(mom[-9].binder, 4),
                                     "note": "sweeps since the hot start, not equilibrated"}}


def workloads(repeats: int = 3) -> dict:
    """The other single-lattice workloads, timed end to end on their own streams (measurements
    included): 2D int8 L=1024 at beta=0.4406868, seed 1, hot, 100 sweeps then (M, E); 2D int8
    L=8192 at T in {1.5, 2.269, 3.0}, 200 sweeps each; 3D int8 L=256 at K_c, seed 7, both
    algorithms, hot, 2000 sweeps with (E, |M|) every 10 sweeps.

    Each configuration runs once untimed on a lattice of its own (the process's first use of a
    kernel or a read-back path costs up to a few milliseconds, as much as the whole L=1024 run),
    then on `repeats` fresh lattices, each clocked up by 20 sweeps before its timed region; the
    median is reported. Every fresh lattice follows the same trajectory, so the reported series
    is that of a lattice 20 sweeps past its hot start."""
    from pyising_amd import _lib
    from pyising_amd.backends.hip_backend import HipInt8
    from pyising_amd.lattice import BETA_C_3D

    out = {}
    stream = _lib.Stream()

    def once(make, sweeps, every):
        sim = make()
        try:
            sim.sweep(20)  # clocks up
            sim.synchronize()
            series = []
            t0 = time.perf_counter()
            if every:  # exact (M, E) gathered on the device, read back once
                series = [tuple(x) for x in sim.measure_series(sweeps // every, every).tolist()]
            else:
                sim.sweep(sweeps)
                series.append(sim.observables_exact())
            dt = time.perf_counter() - t0
            return sweeps * sim.n_spins / dt / 1e9, dt, series
        finally:
            sim.close()

    def timed(make, sweeps, every=None):
        once(make, sweeps, every)  # untimed
        runs = sorted((once(make, sweeps, every) for _ in range(repeats)), key=lambda r: r[1])
        return runs[len(runs) // 2]

    rate, dt, ser = timed(lambda: HipInt8((1024, 1024), 0.4406868, seed=1, stream=stream.ptr), 100)
    out["int8_2d_L1024_100_sweeps"] = {"flips_per_ns": round(rate, 1), "ms": round(dt * 1e3, 3),
                                       "M_E_after": list(ser[-1])}
    for T in (1.5, 2.269, 3.0):
        rate, dt, ser = timed(lambda: HipInt8((8192, 8192), 1.0 / T, seed=1, stream=stream.ptr), 200)
        out[f"int8_2d_L8192_T{T}"] = {"flips_per_ns": round(rate, 1), "M_E_after": list(ser[-1])}
    for alg in ("metropolis", "heatbath"):
        rate, dt, ser = timed(lambda: HipInt8((256, 256, 256), BETA_C_3D, seed=7, algorithm=alg,
                                              stream=stream.ptr), 2000, every=10)
        out[f"int8_3d_L256_Kc_{alg}_2000_sweeps_obs_every_10"] = {
            "flips_per_ns": round(rate, 1), "ms": round(dt * 1e3, 1),
            "mean_abs_m": round(float(np.mean([abs(m) for m, _ in ser[len(ser) // 2:]])) / 256**3, 5),
            "mean_e": round(float(np.mean([e for _, e in ser[len(ser) // 2:]])) / 256**3, 5)}
    stream.close()
    out["timing"] = f"one untimed run, then the median of {repeats} fresh lattices"
    return out


def serving(n_requests: int = 300, rate: float = 200.0, seed: int = 0) -> dict:
    """The serving loop on the Poisson request stream: arrivals at ``rate`` per second
    (2D L = 32..4096 int8/msc with p = 0.7, else 3D L = 16..256 int8 Metropolis/heat bath;
    T in [0.8, 1.2] T_c; sweeps log-uniform in [100, 10^4]; hot starts), served in real time
    with continuous batching. Latency = finish - arrival; aggregate flips/ns over the wall time
    of the whole stream (requests arrive over ~1.5 s)."""
    from pyising_amd.serve import Server, poisson_requests

    reqs = poisson_requests(n_requests, rate=rate, seed=seed)
    work = sum(r.sweeps * r.n_spins for r in reqs)
    srv = Server(budget_bytes=8 << 30, max_slots=256, step_sweeps=25, max_spins=1 << 24)
    rep = srv.run(reqs, realtime=True, timeout=600)
    srv.close()
    out = rep.summary()
    out.update({"rate_per_s": rate, "arrival_span_s": round(reqs[-1].arrival, 3),
                "offered_flips_per_ns": round(work / reqs[-1].arrival / 1e9, 1),
                "stream": f"poisson_requests(n={n_requests}, rate={rate}, seed={seed})"})
    return out


def settle_clocks(sim, e0, e1, min_ms: float, max_ms: float = 3000.0, chunk: int = 100,
                  window: int = 8, tol: float = 0.01) -> tuple[int, float]:
    """Sweep ``sim`` in chunks until at least ``min_ms`` of GPU time has passed and the last
    ``window`` chunks agree within ``tol`` (the clocks have stopped moving), or ``max_ms`` has
    passed (events on the simulation's stream); returns (sweeps run, GPU milliseconds). A GPU
    that has idled through the set-up takes a few hundred milliseconds of sustained load to
    reach its steady clock, and loses it again within about 10 ms of idling."""
    sweeps, ms, recent = 0, 0.0, []
    while ms < max_ms:
        e0.record(sim.stream_ptr)
        sim.sweep(chunk)
        e1.record(sim.stream_ptr)
        e1.sync()
        t = e0.elapsed_ms(e1)
        ms += t
        sweeps += chunk
        recent = (recent + [t])[-window:]
        if ms >= min_ms and len(recent) == window and max(recent) <= (1 + tol) * min(recent):
            break
    return sweeps, ms


def traffic_from_profile() -> tuple[int | None, str]:
    """HBM bytes per msc half-sweep launch from the committed rocprofv3 PMC summary, only if
    that summary was taken on the kernel sources built now (their digest is recorded in it)."""
    from pyising_amd.probes import kernel_source_digest

    try:
        d = json.loads(PMC_FILE.read_text())
        recorded, value = d.get("kernel_sources_sha16"), int(d["hbm_bytes_per_launch"])
    except (OSError, KeyError, ValueError):
        return None, f"no PMC summary at {PMC_FILE.name}"
    now = kernel_source_digest()
    if recorded != now:
        return None, (f"{PMC_FILE.name} was taken on kernel sources {recorded}, "
                      f"built sources are {now}: not used")
    return value, f"{PMC_FILE.name} (FETCH_SIZE + WRITE_SIZE, kernel sources {now})"


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=1000)
    ap.add_argument("--settle-ms", type=float, default=500.0,
                    help="GPU milliseconds of sweeps (at least) before --warmup, to settle the "
                         "clocks; continues until the sweep rate is steady (at most 3 s)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-ceilings", action="store_true")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--no-workloads", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (got --gpus {args.gpus}); sharding a lattice "
              "or a batch across GPUs is designed but not implemented", file=sys.stderr)
        return 2
    if int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py is single-process; WORLD_SIZE > 1 is not supported", file=sys.stderr)
        return 2

    import ctypes as C

    from pyising_amd import _lib
    from pyising_amd.backends.hip_backend import HipMultiSpin
    from pyising_amd.lattice import BETA_C_2D

    lib = _lib.lib()
    name = C.create_string_buffer(64)
    n_cu = C.c_int()
    total = C.c_longlong()
    _lib.check(lib.ising_device_info(name, 64, C.byref(n_cu), C.byref(total)))

    shape = (L_HEADLINE, L_HEADLINE)
    n_spins = L_HEADLINE * L_HEADLINE
    # the simulation's own non-blocking stream (the `stream=` keyword); the legacy null stream
    # measures 2-4 % lower for the same launches
    stream = _lib.Stream()
    e0, e1 = _lib.Event(), _lib.Event()
    sim = HipMultiSpin(shape, BETA_C_2D, seed=1, init="hot", stream=stream.ptr)
    # clock settle (outside --warmup and the timed region): the GPU's clocks need tens of
    # milliseconds of sustained load to reach their steady level after the idle set-up
    settle_sweeps, settle_ms = settle_clocks(sim, e0, e1, args.settle_ms)
    sim.sweep(args.warmup)
    _lib.device_sync()  # (the fused sweeps' status is checked with the observables below)
    gc.disable()  # no collection pass inside the (sub-millisecond at --steps 20) timed region
    t0 = time.perf_counter()
    e0.record(sim.stream_ptr)
    sim.sweep(args.steps)
    e1.record(sim.stream_ptr)
    while not e1.done():  # poll, so the end is seen without the runtime's blocking wake-up
        pass
    _lib.device_sync()
    wall = time.perf_counter() - t0
    gc.enable()
    gpu_ms = e0.elapsed_ms(e1)
    flips_per_ns = args.steps * n_spins / wall / 1e9
    launches = 2 * args.steps
    kernel_s = gpu_ms * 1e-3 / launches  # average half-sweep launch, events on its stream
    algo_bytes = BYTES_PER_UPDATE * n_spins / 2  # per half-sweep launch
    achieved = algo_bytes / kernel_s / 1e9
    m, e = sim.observables_exact()
    stats = sim.philox_stats(2)
    traffic, traffic_note = traffic_from_profile()

    result = {
        "metric": "spin flips/ns (2D msc Metropolis, L=16384, T=T_c)",
        "value": round(flips_per_ns, 2),
        "unit": "flips/ns",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall * 1e3 / args.steps, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u64",
        "data": "synthetic (hot start: numpy default_rng(1) draws)",
        "config": {"workload": "2D multi-spin-coded Metropolis sweep, L=16384 torus, "
                               "beta=1/T_c, seed=1, hot start; one step = one sweep",
                   "L": L_HEADLINE, "beta": BETA_C_2D, "seed": 1, "kernel": sim.kernel()},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": traffic, "traffic_source": traffic_note,
                     "algorithmic_bytes_per_launch": int(algo_bytes),
                     "avg_launch_us": round(kernel_s * 1e6, 3)},
        "clock_settle": {"sweeps": settle_sweeps, "gpu_ms": round(settle_ms, 1),
                         "note": "sweeps run before --warmup, outside the timed region"},
        "device": {"arch": name.value.decode(), "cu": n_cu.value,
                   "hbm_GiB": round(total.value / 2**30, 1)},
        "final_M": m, "final_E": e,
        "philox_calls_per_word": round(stats["calls_per_word"], 4),
        "philox_slots_per_word": round(stats["slots_per_word"], 4),
    }
    sim.close()
    for h in (e0, e1, stream):
        h.close()
    del sim
    if not args.no_ceilings:
        from pyising_amd.probes import copy_bandwidth, philox_rate

        bw = copy_bandwidth()
        ph = philox_rate()
        result["ceilings"] = {
            "copy_bandwidth_GBs": round(bw, 1),
            "pct_copy_bw": round(100 * flips_per_ns * BYTES_PER_UPDATE / bw, 2),
            "philox_calls_per_ns": round(ph, 1),
            "rng_ceiling_flips_per_ns": round(ph * 64 / stats["calls_per_word"], 1),
            "memory_ceiling_flips_per_ns": round(bw / BYTES_PER_UPDATE, 1),
        }
    if not args.no_batch:
        result["batched_scan"] = batch_vs_plain()
    if not args.no_workloads:
        result["workloads"] = workloads()
    if not args.no_serve:
        result["serving"] = serving()
    if not args.no_cpu_baseline:
        result["cpu_baseline"] = cpu_baseline()
    print(json.dumps(result), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
