"""Command line: ``python -m pyising_amd list | run | bench | validate``.

The commands, their options and their output lines are those of the reference's command line
(src/pyising/cli.py:22-110: ``--backend --variant --dim --L --T --algorithm --seed --init
--sweeps --warmup``, ``--repeats`` for bench, ``--quick --out`` for validate), with ``hip`` as
the backend; the reference's ``movie`` command is not provided. Timings bracket the sweeps with
device synchronisation, as the reference's do.
"""

from __future__ import annotations

import argparse
import statistics
import sys
import time
from dataclasses import dataclass

from . import backends as B
from .lattice import TC_2D, TC_3D

#: options shared by the simulation commands: flag -> (type, default, extra argparse keywords)
SIM_OPTIONS: dict[str, tuple] = {
    "--backend": (str, "hip", {"choices": B.BACKENDS}),
    "--variant": (str, "int8", {}),
    "--dim": (int, 2, {"choices": (2, 3)}),
    "--L": (int, 1024, {}),
    "--T": (float, None, {"help": "temperature (default: the critical temperature of --dim)"}),
    "--algorithm": (str, "metropolis", {"choices": ("metropolis", "heatbath")}),
    "--seed": (int, 0, {}),
    "--init": (str, "hot", {"choices": ("hot", "cold")}),
    "--sweeps": (int, 100, {}),
    "--warmup": (int, 10, {}),
}


@dataclass(frozen=True)
class Workload:
    """One simulation as the options describe it."""

    backend: str
    variant: str
    shape: tuple
    temperature: float
    seed: int
    algorithm: str
    init: str

    @classmethod
    def from_args(cls, a: argparse.Namespace) -> "Workload":
        critical = TC_2D if a.dim == 2 else TC_3D
        return cls(a.backend, a.variant, (a.L,) * a.dim,
                   critical if a.T is None else a.T, a.seed, a.algorithm, a.init)

    def simulation(self):
        return B.create(self.backend, self.variant, shape=self.shape, beta=1.0 / self.temperature,
                        seed=self.seed, algorithm=self.algorithm, init=self.init)


def _timed(sim, sweeps: int) -> float:
    """Seconds for ``sweeps`` sweeps, the device synchronised before and after."""
    sim.synchronize()
    start = time.perf_counter()
    sim.sweep(sweeps)
    sim.synchronize()
    return time.perf_counter() - start


def _flips_per_ns(sim, sweeps: int, seconds: float) -> float:
    return sweeps * sim.n_spins / seconds * 1e-9


# ------------------------------------------------------------------------- commands ----
def do_list(a: argparse.Namespace) -> int:
    """Every registry entry, whether it can run here, and what it supports."""
    for backend, variant in B.REGISTRY:
        cls = B.load(backend, variant)
        if cls is None:
            status, detail = "unavailable", ""
        else:
            status = "ok"
            detail = (f"dims={cls.dims} algs={cls.algorithms} "
                      f"B/update={cls.bytes_per_update:g}")
        print(f"{backend:>14} / {variant:<18} {status:<12} {detail}")
    return 0


def do_run(a: argparse.Namespace) -> int:
    """One timed run: m, e after the sweeps and the rate of this single run."""
    w = Workload.from_args(a)
    sim = w.simulation()
    sim.sweep(a.warmup)
    seconds = _timed(sim, a.sweeps)
    m, e = sim.observables()
    print(f"{sim.backend}/{sim.variant} {sim.algorithm} shape={sim.shape} T={1 / sim.beta:.5f}")
    print(f"m = {m:+.6f}   e = {e:.6f}   {_flips_per_ns(sim, a.sweeps, seconds):.2f} flips/ns "
          f"({seconds:.3f} s, single run)")
    return 0


def do_bench(a: argparse.Namespace) -> int:
    """Median of --repeats timed runs of --sweeps sweeps (one configuration)."""
    w = Workload.from_args(a)
    sim = w.simulation()
    sim.sweep(a.warmup)
    rates = [_flips_per_ns(sim, a.sweeps, _timed(sim, a.sweeps)) for _ in range(a.repeats)]
    print(f"{sim.backend}/{sim.variant} L={a.L} dim={a.dim}: median "
          f"{statistics.median(rates):.2f} flips/ns (min {min(rates):.2f}, max {max(rates):.2f});"
          f" use python -m pyising_amd.report for the full size sweep")
    return 0


def do_validate(a: argparse.Namespace) -> int:
    """The physics validation suite; exit status 1 unless every backend run was identical."""
    from .validate import run

    summary = run(quick=a.quick, out=a.out)["summary"]
    width = max(len(k) for k in summary)
    for key, value in summary.items():
        print(f"{key:>{max(40, width)}}: {value}")
    return 0 if summary["all_identical"] else 1


#: name -> (help, handler, extra options beyond SIM_OPTIONS, takes SIM_OPTIONS)
COMMANDS: dict[str, tuple] = {
    "list": ("list backends/variants and availability", do_list, {}, False),
    "run": ("run a simulation and print m, e, throughput", do_run, {}, True),
    "bench": ("quick timing of one configuration", do_bench,
              {"--repeats": (int, 5, {})}, True),
    "validate": ("physics validation suite", do_validate,
                 {"--quick": (None, False, {"action": "store_true"}),
                  "--out": (str, "results/validation.json", {})}, False),
}


def _add(sp: argparse.ArgumentParser, flag: str, spec: tuple) -> None:
    typ, default, extra = spec
    kw = dict(extra)
    if typ is not None:
        kw["type"] = typ
    sp.add_argument(flag, default=default, **kw)


def parser() -> argparse.ArgumentParser:
    p = argparse.ArgumentParser(prog="pyising_amd", description=__doc__,
                                formatter_class=argparse.RawDescriptionHelpFormatter)
    sub = p.add_subparsers(dest="cmd", required=True)
    for name, (help_text, handler, extra, with_sim) in COMMANDS.items():
        sp = sub.add_parser(name, help=help_text)
        sp.set_defaults(handler=handler)
        for flag, spec in (SIM_OPTIONS.items() if with_sim else ()):
            _add(sp, flag, spec)
        for flag, spec in extra.items():
            _add(sp, flag, spec)
    return p


def main(argv: list[str] | None = None) -> int:
    args = parser().parse_args(argv)
    return args.handler(args)


if __name__ == "__main__":
    sys.exit(main())
