"""Benchmark report of the hip backend in the reference's results layout
(bench/run.py:208-240: top-level machine / ceiling / provisional / mode / runs / seconds, one
run record per backend, variant, algorithm and size), plus a roofline figure and a markdown
table (figures/make_figures.py:127-155).

    python -m pyising_amd.report --quick            # ~1 min -> results/mi355x-quick.json
    python -m pyising_amd.report --full             # full size plan -> results/mi355x.json
    python -m pyising_amd.report --dev              # full plan, lattices < 600 MB

Metric: attempted spin updates per ns at T_c, hot start. Traffic model per update: own spin read
and written, opposite-colour neighbours read once = 3 spin-sized transfers (3 B int8, 3/8 B
packed). Effective bandwidth = flips/ns x bytes per update; the roofline is the lower of the
copy-bandwidth ceiling and the Philox ceiling (Philox calls per ns / calls per update).
"""

from __future__ import annotations

import argparse
import datetime as dt
import gc
import json
import math
import platform
import statistics
import sys
import time
from pathlib import Path

import numpy as np

from . import backends as B
from .lattice import BETA_C_2D, BETA_C_3D

HBM_SPEC_GBS = 8000.0  # MI355X HBM3E, published
FP32_TFLOPS = 157.3  # MI355X dense fp32 (not used by the integer kernels)

FULL_2D = [1024, 2048, 4096, 8192, 16384, 32768, 65536]
FULL_MSC = [1024, 2048, 4096, 8192, 16384, 32768, 65536, 131072]
FULL_3D = [64, 128, 256, 512, 1024]
#: variant, dim, sizes, algorithm (the hip rows of bench/run.py:39-62; "max" = largest lattice
#: whose set-up fits in 80 % of free HBM)
PLAN_FULL: list[tuple[str, int, list, str]] = [
    ("int8", 2, FULL_2D + ["max"], "metropolis"),
    ("msc", 2, FULL_MSC + ["max"], "metropolis"),
    ("msc-eager32", 2, [4096, 16384, 65536], "metropolis"),
    ("int8", 2, [16384], "heatbath"),
    ("int8", 3, FULL_3D + ["max"], "metropolis"),
    ("int8", 3, [256], "heatbath"),
]
PLAN_QUICK: list[tuple[str, int, list, str]] = [
    ("int8", 2, [1024, 4096], "metropolis"),
    ("msc", 2, [1024, 8192], "metropolis"),
    ("msc-eager32", 2, [8192], "metropolis"),
    ("int8", 3, [128], "metropolis"),
]


def machine_info() -> dict:
    import torch

    from .probes import device_info

    d = device_info()
    info = {"gpu": d["arch"], "compute_units": d["cu"], "hbm_GiB": round(d["hbm_bytes"] / 2**30, 1),
            "python": platform.python_version(), "torch": torch.__version__,
            "numpy": np.__version__,
            "date": dt.datetime.now(dt.timezone.utc).isoformat(timespec="seconds")}
    try:
        info["rocm"] = Path("/opt/rocm/.info/version").read_text().strip()
    except OSError:
        info["rocm"] = None
    return info


def msc_calls_per_word(L: int = 4096, mode: int | None = None, sweeps: int = 20) -> dict:
    """Philox calls the lazy comparison needs per packed word at T_c (mean), and the lane slots
    the SIMDs execute per word with the given kernel mode (bench/roofline.py:59-78 counts the
    same two quantities: mean and per-warp maximum)."""
    from .backends.hip_backend import HipMultiSpin

    cls = HipMultiSpin if mode is None else type("M", (HipMultiSpin,), {"mode": mode})
    sim = cls((L, L), BETA_C_2D, seed=1, init="hot")
    sim.sweep(sweeps)
    return sim.philox_stats(4)


def measure_ceilings(quick: bool = False) -> dict:
    """Copy bandwidth, Philox rate and msc Philox work (bench/roofline.py:97-117 keys)."""
    from .backends.hip_backend import MSC_COMPACT, MSC_LOCKSTEP
    from .probes import copy_bandwidth, philox_rate

    nbytes = (1 << 28) if quick else (1 << 32)
    bw = copy_bandwidth(nbytes, repeats=10 if quick else 20)
    ph = philox_rate(128 if quick else 512)
    L = 2048 if quick else 4096
    queue = msc_calls_per_word(L, MSC_COMPACT)
    lockstep = msc_calls_per_word(L, MSC_LOCKSTEP)
    return {
        "copy_bandwidth_GBs": round(bw, 1),
        "copy_bandwidth_measured_GBs": round(bw, 1),
        "copy_bandwidth_source": "measured",
        "spec_bandwidth_GBs": HBM_SPEC_GBS,
        "fp32_tflops": FP32_TFLOPS,
        "philox_calls_per_ns": round(ph, 2),
        "msc_philox_calls_per_word_at_Tc": round(queue["calls_per_word"], 3),
        # lockstep waves run every lane for the wave's largest count, as the reference's warps
        "msc_philox_calls_per_word_warp_max_at_Tc": round(lockstep["slots_per_word"], 3),
        "msc_philox_slots_per_word_queue_at_Tc": round(queue["slots_per_word"], 3),
        "copy_bandwidth_bytes": nbytes,
    }


def max_size(cls, dim: int, frac: float = 0.8) -> int:
    """Largest lattice whose set-up fits in ``frac`` of the free device memory: the state
    (bytes_per_update / 3 per spin) plus the one-byte-per-spin full lattice the hot start
    writes before the split into colours."""
    from . import _lib

    free = _lib.mem_info()[0]
    per_spin = cls.bytes_per_update / 3.0 + 1.0
    L = int((frac * free / per_spin) ** (1.0 / dim))
    # int8: Lx % 32 == 0 gives 16 sites per lane; the kernels' 32-bit chunk index (sites / 32)
    # or word index (spins / 128 per colour) must stay 2^20 below 2^32
    m = cls.layout_multiple if cls.stream == "msc" else 32
    per_index = 128 if cls.stream == "msc" else 32
    L = max(m, (L // m) * m)
    while L**dim // per_index + (1 << 20) > (1 << 32) - 1:
        L -= m
    return L


def time_one(cls, shape, algorithm: str, beta: float, repeats: int, min_time: float,
             warm_time: float = 0.25) -> dict:
    """Size the run to >= ``min_time`` s per repeat and return median/min/max flips/ns
    (bench/run.py:121-142). Before timing, ``warm_time`` s of sweeps at the timed rate lift the
    GPU out of its idle clocks (they ramp over some tens of ms)."""
    sim = cls(shape, beta, seed=12345, algorithm=algorithm, init="hot")
    sim.sweep(2)
    sim.synchronize()
    t0 = time.perf_counter()
    sim.sweep(1)
    sim.synchronize()
    t1 = max(time.perf_counter() - t0, 1e-6)
    sim.sweep(max(1, math.ceil(warm_time / t1)))
    sim.synchronize()
    t0 = time.perf_counter()
    sim.sweep(10)
    sim.synchronize()
    t1 = max((time.perf_counter() - t0) / 10, 1e-7)
    n = max(1, math.ceil(min_time / t1))
    rates = []
    for _ in range(repeats):
        t0 = time.perf_counter()
        sim.sweep(n)
        sim.synchronize()
        rates.append(n * sim.n_spins / (time.perf_counter() - t0) / 1e9)
    m, e = sim.observables()
    del sim
    return {"median": statistics.median(rates), "min": min(rates), "max": max(rates),
            "sweeps_per_repeat": n, "m": m, "e": e}


def calls_per_update(cls, ceiling: dict) -> float:
    """Philox calls per attempted update: 1 call per 4 int8 sites; for msc the measured lane
    slots per 64-spin word of the default kernel; 16 per word for the eager ablation."""
    if cls.stream != "msc":
        return 0.25
    if cls.variant == "msc-eager32":
        return 16.0 / 64.0
    return ceiling["msc_philox_slots_per_word_queue_at_Tc"] / 64.0


def run_record(cls, variant: str, algorithm: str, shape, r: dict, ceiling: dict,
               repeats: int, beta: float) -> dict:
    """One run in the layout of bench/run.py:196-217."""
    bw = ceiling["copy_bandwidth_GBs"]
    bpu = cls.bytes_per_update
    eff = r["median"] * bpu
    cpu = calls_per_update(cls, ceiling)
    rng = ceiling["philox_calls_per_ns"] / cpu
    return {
        "backend": "hip", "variant": variant, "algorithm": algorithm,
        "size": list(shape), "dim": len(shape),
        "dtype": "uint64-packed-1bit" if cls.stream == "msc" else "int8",
        "metric": "flips/ns", "value": round(r["median"], 3), "repeats": repeats,
        "median": round(r["median"], 3), "min": round(r["min"], 3), "max": round(r["max"], 3),
        "sweeps_per_repeat": r["sweeps_per_repeat"],
        "bytes_per_unit": bpu, "effective_GBs": round(eff, 2),
        "pct_copy_bw": round(100 * eff / bw, 2),
        "pct_spec_bw": round(100 * eff / ceiling["spec_bandwidth_GBs"], 2),
        "flops_per_unit": None, "pct_peak_flops": None,
        "memory_ceiling_flips_per_ns": round(bw / bpu, 1),
        "philox_calls_per_update": round(cpu, 5),
        "rng_ceiling_flips_per_ns": round(rng, 1),
        "pct_roofline": round(100 * r["median"] / min(bw / bpu, rng), 2),
        "temperature": 1.0 / beta, "final_m": r["m"], "final_e": r["e"],
        "notes": f"{algorithm}, T=Tc, hot start",
    }


def run(mode: str = "quick", out: str | Path | None = None, repeats: int | None = None,
        only: set[str] | None = None, log=print) -> dict:
    """Measure the ceilings and every plan entry; returns (and writes) the report dict."""
    from . import _lib

    plan = PLAN_QUICK if mode == "quick" else PLAN_FULL
    min_time = 0.1 if mode == "quick" else 0.5
    repeats = repeats or (3 if mode == "quick" else 5)
    t_start = time.time()
    log("measuring ceilings ...")
    ceiling = measure_ceilings(quick=mode != "full")
    log(json.dumps(ceiling))
    runs = []
    for variant, dim, sizes, algorithm in plan:
        if only and not ({"hip", variant} & only):
            continue
        cls = B.load("hip", variant)
        if cls is None:
            log(f"skip hip/{variant}: unavailable")
            continue
        beta = BETA_C_2D if dim == 2 else BETA_C_3D
        for L in sizes:
            if L == "max" and mode == "dev":
                continue
            L = max_size(cls, dim) if L == "max" else int(L)
            shape = (L,) * dim
            if mode == "dev" and L**dim * (cls.bytes_per_update / 3.0 + 1.0) > 600e6:
                continue
            try:
                r = time_one(cls, shape, algorithm, beta, repeats, min_time)
            except (RuntimeError, MemoryError, ValueError) as exc:  # e.g. out of memory
                log(f"FAIL hip/{variant} {shape}: {type(exc).__name__}: {exc}"[:300])
                gc.collect()
                continue
            finally:
                gc.collect()  # release the lattice's device memory before the next size
            rec = run_record(cls, variant, algorithm, shape, r, ceiling, repeats, beta)
            runs.append(rec)
            log(f"{'hip':>6}/{variant:<12} {algorithm[:5]} {str(shape):>22}: "
                f"{rec['median']:9.2f} flips/ns [{rec['min']:.2f}, {rec['max']:.2f}] "
                f"{rec['pct_copy_bw']:6.1f}% copy BW {rec['pct_roofline']:6.1f}% roofline "
                f"({time.time() - t_start:5.0f} s)")
        _lib.device_sync()
    res = {"machine": machine_info(), "ceiling": ceiling, "provisional": mode != "full",
           "mode": mode, "runs": runs, "seconds": round(time.time() - t_start, 1)}
    if out is not None:
        out = Path(out)
        out.parent.mkdir(parents=True, exist_ok=True)
        out.write_text(json.dumps(res, indent=1))
        log(f"wrote {out} in {res['seconds']} s")
    return res


def best_per_variant(report: dict) -> dict:
    """Fastest 2D Metropolis run of each variant (figures/make_figures.py best_per_variant)."""
    best: dict = {}
    for r in report["runs"]:
        if r["dim"] != 2 or r["algorithm"] != "metropolis":
            continue
        k = (r["backend"], r["variant"])
        if k not in best or r["median"] > best[k]["median"]:
            best[k] = r
    return best


def markdown(report: dict) -> str:
    """The runs as a markdown table with their roofline fractions."""
    c = report["ceiling"]
    lines = [f"Ceilings: copy {c['copy_bandwidth_GBs']:.0f} GB/s (spec {c['spec_bandwidth_GBs']:.0f}),"
             f" Philox {c['philox_calls_per_ns']:.0f} calls/ns, msc Philox calls/word at T_c "
             f"{c['msc_philox_calls_per_word_at_Tc']:.3f} (lockstep slots "
             f"{c['msc_philox_calls_per_word_warp_max_at_Tc']:.3f}).", "",
             "| variant | alg | size | flips/ns (median) | min..max | GB/s eff. | % copy BW "
             "| RNG ceiling | % roofline | final m | final e |",
             "|---|---|---|---|---|---|---|---|---|---|---|"]
    for r in report["runs"]:
        lines.append(f"| {r['variant']} | {r['algorithm']} | {'x'.join(map(str, r['size']))} "
                     f"| {r['median']:.2f} | {r['min']:.2f}..{r['max']:.2f} "
                     f"| {r['effective_GBs']:.0f} | {r['pct_copy_bw']:.1f} "
                     f"| {r['rng_ceiling_flips_per_ns']:.0f} | {r['pct_roofline']:.1f} "
                     f"| {r['final_m']:+.4f} | {r['final_e']:.4f} |")
    return "\n".join(lines) + "\n"


def roofline_figure(report: dict, path: str | Path) -> None:
    """Flips/ns against spin updates per DRAM byte with the copy-bandwidth slope, the spec
    slope and the two Philox roofs (figures/make_figures.py:127-155)."""
    import matplotlib

    matplotlib.use("Agg")
    import matplotlib.pyplot as plt

    c = report["ceiling"]
    fig, ax = plt.subplots(figsize=(4.5, 3.2))
    x = np.logspace(-1.2, 1.3, 100)
    ax.plot(x, c["copy_bandwidth_GBs"] * x, color="k", lw=1.2)
    ax.plot(x, c["spec_bandwidth_GBs"] * x, color="0.6", lw=0.8, ls="--")
    ax.text(1.3, c["copy_bandwidth_GBs"] * 1.3 * 0.25, f"copy {c['copy_bandwidth_GBs']:.0f} GB/s",
            rotation=33, fontsize=6.5, ha="center")
    ph = c["philox_calls_per_ns"]
    roofs = ((0.25, "Philox roof, 4 sites/call", (0.06, 0.6)),
             (c["msc_philox_slots_per_word_queue_at_Tc"] / 64, "Philox roof, lazy bits (queue)",
              (0.6, 20)))
    for cps, name, xr in roofs:
        ax.hlines(ph / cps, *xr, color="0.4", lw=0.9, ls=":")
        ax.text(xr[0] * 1.05, ph / cps * 1.1, name, fontsize=6.5, color="0.3")
    colours = {"int8": "tab:blue", "msc": "tab:red", "msc-eager32": "tab:orange"}
    for (_, variant), r in best_per_variant(report).items():
        ax.plot(1 / r["bytes_per_unit"], r["median"], "o", ms=5,
                color=colours.get(variant, "tab:gray"), label=f"hip/{variant}")
    ax.legend(fontsize=6.5, frameon=False, loc="lower right")
    ax.set_xscale("log")
    ax.set_yscale("log")
    ax.set_xlabel("spin updates per DRAM byte")
    ax.set_ylabel("flips / ns")
    ax.set_title(report["machine"].get("gpu", ""), fontsize=8)
    fig.tight_layout()
    Path(path).parent.mkdir(parents=True, exist_ok=True)
    fig.savefig(path, dpi=150)
    plt.close(fig)


def scaling_figure(report: dict, path: str | Path) -> None:
    """Flips/ns against the number of spins for every variant, 2D and 3D side by side, with the
    copy-bandwidth ceilings (figures/make_figures.py:99-124)."""
    import matplotlib

    matplotlib.use("Agg")
    import matplotlib.pyplot as plt

    c = report["ceiling"]
    fig, axes = plt.subplots(1, 2, figsize=(7.0, 2.9), sharey=True)
    for ax, dim in ((axes[0], 2), (axes[1], 3)):
        series: dict = {}
        for r in report["runs"]:
            if r["dim"] == dim:
                series.setdefault((r["variant"], r["algorithm"]), []).append(r)
        for (var, alg), rs in sorted(series.items()):
            rs.sort(key=lambda r: r["size"][0])
            ls = ":" if "eager" in var else ("--" if alg == "heatbath" else "-")
            ax.plot([float(np.prod(r["size"])) for r in rs], [r["median"] for r in rs], ls,
                    marker="o", ms=3, lw=1.2, label=f"hip/{var} {alg}")
        for bpu, name in ((3.0, "int8"), (3.0 / 8.0, "msc")):
            if dim == 3 and name == "msc":
                continue
            ax.axhline(c["copy_bandwidth_GBs"] / bpu, color="k", lw=0.8)
            ax.text(0.02, c["copy_bandwidth_GBs"] / bpu * 1.08, f"copy ceiling, {name}",
                    fontsize=6.5, transform=ax.get_yaxis_transform())
        ax.set_xscale("log")
        ax.set_yscale("log")
        ax.set_xlabel(f"spins $N = L^{dim}$")
        ax.set_title(f"{dim}D", fontsize=9)
        if series:
            ax.legend(fontsize=6, frameon=False, loc="lower right")
    axes[0].set_ylabel("flips / ns (median)")
    fig.tight_layout()
    Path(path).parent.mkdir(parents=True, exist_ok=True)
    fig.savefig(path, dpi=150)
    plt.close(fig)


def main(argv: list[str] | None = None) -> int:
    ap = argparse.ArgumentParser(prog="python -m pyising_amd.report", description=__doc__,
                                 formatter_class=argparse.RawTextHelpFormatter)
    g = ap.add_mutually_exclusive_group(required=True)
    g.add_argument("--quick", action="store_true")
    g.add_argument("--full", action="store_true")
    g.add_argument("--dev", action="store_true", help="full plan, lattices < 600 MB")
    ap.add_argument("--out", default=None)
    ap.add_argument("--repeats", type=int, default=None)
    ap.add_argument("--only", default=None, help="comma list of variants (or 'hip')")
    ap.add_argument("--markdown", default=None, help="also write the runs as a markdown table")
    ap.add_argument("--figure", default=None, help="also draw the roofline figure (png/pdf)")
    ap.add_argument("--scaling-figure", default=None, help="also draw flips/ns against size")
    ap.add_argument("--from-json", default=None,
                    help="redraw the table and figures from an existing report (no GPU)")
    args = ap.parse_args(argv)
    mode = "quick" if args.quick else ("dev" if args.dev else "full")
    name = {"quick": "mi355x-quick.json", "dev": "provisional.json", "full": "mi355x.json"}[mode]
    out = Path(args.out) if args.out else Path("results") / name
    only = set(args.only.split(",")) if args.only else None
    if args.from_json:
        res = json.loads(Path(args.from_json).read_text())
    else:
        res = run(mode, out, args.repeats, only, log=lambda s: print(s, flush=True))
    if args.markdown:
        Path(args.markdown).write_text(markdown(res))
    if args.figure:
        roofline_figure(res, args.figure)
    if args.scaling_figure:
        scaling_figure(res, args.scaling_figure)
    return 0


if __name__ == "__main__":
    sys.exit(main())
