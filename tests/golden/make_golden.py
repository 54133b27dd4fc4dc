"""Writes the golden fixtures of this directory (run from the repository root on a CPU):

    python tests/golden/make_golden.py            (known answers and end states)
    python tests/golden/make_golden.py --long     (full-length runs of the configurations)

* ``reference_known_answers.json`` -- the known-answer data the reference's own tests hold,
  each entry citing the test it comes from: the Random123 Philox4x32-10 vectors
  (tests/test_philox.py:6-12), the seed split and the int8 stream layout
  (tests/test_philox.py:29-34), threshold-table entries (tests/test_lattice.py:33-38: the
  test states them as floor(exp(-2) 2^32) and floor(exp(-4) 2^32); they are stored here as the
  integers those expressions give), the cold-start observables (tests/test_backends.py:73-77,
  plus the 3D value of the same definition) and the exact 2D constants
  (tests/test_exact.py:30-34).
* ``oracle_end_states.json`` -- end states of the oracle (oracle/ising_oracle.py, and its C
  restatement for the large cases) for the reference's identity-suite cases
  (tests/test_backends.py:22-67), config-shaped small cases, and config-sized cases kept as a
  sha256 of the lattice plus exact (M, E). The lattice digest is sha256 over the full int8
  lattice (+1 / -1) in C order, the layout ``get_lattice()`` returns.

* ``oracle_long_runs.json`` -- the configured workloads at their full lengths, from the C
  restatement: msc L=16384 at T_c (seed 1, hot start) for 1000 sweeps and int8 L=8192 at
  T = 1.5, 2.269, 3.0 for 1000 sweeps, with checkpoints (sha256 + exact M, E) along the way;
  3D int8 L=256 (seed 7) for 2000 sweeps, Metropolis and heat bath, hot and cold starts at K_c
  and hot starts at K = 0.20, 0.24, with the exact (M, E) series every 10 sweeps and the final
  sha256.

The oracle is pinned independently of these files (tests/test_oracle_*.py); the fixtures let
the GPU tests check end states without running the oracle, and catch any later change of the
oracle itself.
"""

from __future__ import annotations

import hashlib
import json
import math
import sys
import time
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parent.parent))

from oracle import c_oracle as CO  # noqa: E402
from oracle import ising_oracle as O  # noqa: E402

BETA_C_2D = 0.5 * math.log(1.0 + math.sqrt(2.0))
BETA_C_3D = 0.221654626


def digest(lat: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(lat, dtype=np.int8).tobytes()).hexdigest()


def known_answers() -> dict:
    def t(beta, ndim, alg):
        return [int(x) for x in O.threshold_table(beta, ndim, alg)]

    return {
        "philox4x32_10": {
            "cites": "tests/test_philox.py:6-12 (Random123 known-answer vectors)",
            "cases": [
                {"counter": [0, 0, 0, 0], "key": [0, 0],
                 "expected": [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]},
                {"counter": [0xFFFFFFFF] * 4, "key": [0xFFFFFFFF] * 2,
                 "expected": [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]},
                {"counter": [0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344],
                 "key": [0xA4093822, 0x299F31D0],
                 "expected": [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]},
            ],
        },
        "split_seed": {
            "cites": "tests/test_philox.py:29-34",
            "cases": [{"seed": 2**40 + 3, "expected": [3, 256]}],
        },
        "site_uniforms_layout": {
            "cites": "tests/test_philox.py:29-34: sites 8..11 of site_uniforms(16, sweep=5, "
                     "color=1, seed=2^40+3) are the four words of Philox((2, 5, 1, 0), key)",
            "n_sites": 16, "sweep": 5, "color": 1, "seed": 2**40 + 3, "quad": 2,
        },
        "threshold_table": {
            "cites": "tests/test_lattice.py:33-38",
            "cases": [
                {"beta": 0.5, "ndim": 2, "algorithm": "metropolis",
                 "entries": {"0": 0xFFFFFFFF, "1": 0xFFFFFFFF, "2": 0xFFFFFFFF,
                             "3": int(np.exp(-2.0) * 2**32), "4": int(np.exp(-4.0) * 2**32)}},
                {"beta": 0.5, "ndim": 3, "algorithm": "heatbath", "length": 7,
                 "entries": {"3": 2**31}},
            ],
            # not from the reference's tests: the tables at the configured betas, for the
            # record (the tests recompute them with 60-digit decimal arithmetic)
            "config_tables": {
                "2d_metropolis_beta_c": t(BETA_C_2D, 2, "metropolis"),
                "2d_metropolis_0.4406868": t(0.4406868, 2, "metropolis"),
                "3d_metropolis_beta_c": t(BETA_C_3D, 3, "metropolis"),
                "3d_heatbath_beta_c": t(BETA_C_3D, 3, "heatbath"),
            },
        },
        "cold_start_observables": {
            "cites": "tests/test_backends.py:73-77 (2D); 3D by the same definition",
            "2d": [1.0, -2.0], "3d": [1.0, -3.0],
        },
        "exact_2d": {
            "cites": "tests/test_exact.py:30-34",
            "tc": 2.269185314213022, "energy_at_tc": -math.sqrt(2.0),
            "magnetization_at_T2": (1.0 - math.sinh(1.0) ** -4) ** 0.125,
            "magnetization_at_T3": 0.0,
        },
    }


def case(name, variant, shape, beta, seed, sweeps, algorithm="metropolis", init="hot",
         init_seed=None, then=None, engine="numpy"):
    """One end-state case. init: "hot", "cold" or "random" (default_rng(init_seed) lattice,
    the reference's random_lattice). then: optional (beta2, sweeps2) run after set_beta."""
    t0 = time.time()
    lat0 = O.random_lattice(tuple(shape), init_seed) if init == "random" else init
    if engine == "numpy":
        cls = O.Int8Oracle if variant == "int8" else O.MscOracle
        sim = cls(tuple(shape), beta, seed=seed, algorithm=algorithm, init=lat0)
    else:
        cls = CO.CInt8 if variant == "int8" else CO.CMsc
        sim = cls(tuple(shape), beta, seed=seed, algorithm=algorithm, init=lat0)
    sim.sweep(sweeps)
    if then is not None:
        sim.set_beta(then[0])
        sim.sweep(then[1])
    lat = sim.get_lattice()
    M, E = (int(x) for x in sim.observables_exact())
    rec = {"name": name, "variant": variant, "shape": list(shape), "beta": beta, "seed": seed,
           "algorithm": algorithm, "init": init, "init_seed": init_seed, "sweeps": sweeps,
           "then": list(then) if then else None, "engine": engine,
           "sha256": digest(lat), "M": M, "E": E}
    print(f"{name}: M={M} E={E} ({time.time() - t0:.1f} s)", flush=True)
    return rec


def end_states() -> list[dict]:
    out = []
    # the reference's identity suite (tests/test_backends.py:22-67)
    for shape in ((16, 24), (8, 8, 16), (64, 32)):
        for alg in ("metropolis", "heatbath"):
            beta = 0.42 if len(shape) == 2 else 0.22
            out.append(case(f"identity_int8_{'x'.join(map(str, shape))}_{alg}", "int8", shape,
                            beta, 2**40 + 7, 6, alg, "random", 5))
    for shape in ((128, 128), (64, 384)):
        out.append(case(f"identity_msc_{'x'.join(map(str, shape))}", "msc", shape, 0.44, 11, 5,
                        init="random", init_seed=9))
    for variant in ("int8", "msc"):
        out.append(case(f"set_beta_{variant}_128x128", variant, (128, 128), 0.3, 3, 2,
                        init="random", init_seed=1, then=(0.5, 3)))
    # config-shaped small cases
    out.append(case("int8_2d_L64_hot_100", "int8", (64, 64), 0.4406868, 1, 100))
    out.append(case("int8_2d_L64_hot_100_heatbath", "int8", (64, 64), 0.4406868, 1, 100,
                    "heatbath"))
    out.append(case("msc_128x256_beta_c_20", "msc", (128, 256), BETA_C_2D, 1, 20))
    for alg in ("metropolis", "heatbath"):
        for K in (0.20, BETA_C_3D, 0.24):
            for init in ("cold", "hot"):
                out.append(case(f"int8_3d_L16_{alg}_K{K}_{init}_40", "int8", (16, 16, 16), K, 7,
                                40, alg, init))
    # config-sized cases (C restatement of the oracle)
    out.append(case("int8_2d_L1024_hot_100", "int8", (1024, 1024), 0.4406868, 1, 100,
                    engine="c"))
    for T in (1.5, 2.269, 3.0):
        out.append(case(f"int8_2d_L2048_T{T}_hot_20", "int8", (2048, 2048), 1.0 / T, 1, 20,
                        engine="c"))
    for T in (1.5, 2.269, 3.0):
        out.append(case(f"int8_2d_L8192_T{T}_hot_10", "int8", (8192, 8192), 1.0 / T, 1, 10,
                        engine="c"))
    for alg in ("metropolis", "heatbath"):
        out.append(case(f"int8_3d_L64_Kc_{alg}_hot_20", "int8", (64, 64, 64), BETA_C_3D, 7, 20,
                        alg, engine="c"))
        out.append(case(f"int8_3d_L256_Kc_{alg}_hot_10", "int8", (256, 256, 256), BETA_C_3D, 7,
                        10, alg, engine="c"))
    out.append(case("msc_L1024_beta_c_hot_50", "msc", (1024, 1024), BETA_C_2D, 1, 50,
                    engine="c"))
    out.append(case("msc_L16384_beta_c_hot_3", "msc", (16384, 16384), BETA_C_2D, 1, 3,
                    engine="c"))
    out.append(case("msc_L32768_beta_c_hot_2", "msc", (32768, 32768), BETA_C_2D, 3, 2,
                    engine="c"))
    for alg in ("metropolis", "heatbath"):
        out.append(case(f"int8_2d_L16384_beta_c_{alg}_hot_2", "int8", (16384, 16384), BETA_C_2D,
                        11, 2, alg, engine="c"))
    # queue-compaction msc in its per-lane (88-word rows) and row-segment (32-word rows)
    # layouts at sizes above the lockstep threshold, and a large 3D int8 lattice
    out.append(case("msc_L11264_beta_c_hot_3", "msc", (11264, 11264), BETA_C_2D, 13, 3,
                    engine="c"))
    out.append(case("msc_16384x4096_beta_c_hot_3", "msc", (16384, 4096), BETA_C_2D, 17, 3,
                    engine="c"))
    out.append(case("int8_3d_L512_Kc_metropolis_hot_2", "int8", (512, 512, 512), BETA_C_3D, 19,
                    2, engine="c"))
    return out


def long_run(name, variant, shape, beta, seed, sweeps, every, algorithm="metropolis",
             init="hot", lattice_every=None):
    """A full-length run of the C restatement: exact (M, E) after every ``every`` sweeps and the
    lattice digest after every ``lattice_every`` sweeps (default: the end only)."""
    t0 = time.time()
    cls = CO.CInt8 if variant == "int8" else CO.CMsc
    sim = cls(tuple(shape), beta, seed=seed, algorithm=algorithm, init=init)
    lattice_every = lattice_every or sweeps
    series, digests = [], {}
    done = 0
    while done < sweeps:
        sim.sweep(every)
        done += every
        series.append([int(x) for x in sim.observables_exact()])
        if done % lattice_every == 0:
            digests[str(done)] = digest(sim.get_lattice())
    print(f"{name}: {sweeps} sweeps, M, E = {series[-1]} ({time.time() - t0:.0f} s)", flush=True)
    return {"name": name, "variant": variant, "shape": list(shape), "beta": beta, "seed": seed,
            "algorithm": algorithm, "init": init, "sweeps": sweeps, "every": every,
            "series": series, "sha256_at": digests}


def long_runs() -> list[dict]:
    out = [long_run("msc_L16384_Tc_hot_1000", "msc", (16384, 16384), BETA_C_2D, 1, 1000, 100,
                    lattice_every=100)]
    for T in (1.5, 2.269, 3.0):
        out.append(long_run(f"int8_2d_L8192_T{T}_hot_1000", "int8", (8192, 8192), 1.0 / T, 1,
                            1000, 100, lattice_every=250))
    for alg in ("metropolis", "heatbath"):
        for K, init in ((BETA_C_3D, "hot"), (BETA_C_3D, "cold"), (0.20, "hot"), (0.24, "hot")):
            out.append(long_run(f"int8_3d_L256_K{K}_{alg}_{init}_2000", "int8", (256, 256, 256),
                                K, 7, 2000, 10, alg, init))
    return out


def main() -> int:
    if "--long" in sys.argv[1:]:
        data = {"numpy": np.__version__, "engine": "oracle C restatement (oracle/ising_oracle.c)",
                "note": "series: exact (M, E) after every `every` sweeps; sha256_at: digest of "
                        "the full int8 lattice (+1/-1, C order) after that many sweeps",
                "cases": long_runs()}
        (HERE / "oracle_long_runs.json").write_text(json.dumps(data) + "\n")
        return 0
    ka = known_answers()
    (HERE / "reference_known_answers.json").write_text(json.dumps(ka, indent=1) + "\n")
    states = {"numpy": np.__version__,
              "note": "sha256 over the full int8 lattice (+1/-1) in C order; M, E exact",
              "cases": end_states()}
    (HERE / "oracle_end_states.json").write_text(json.dumps(states, indent=1) + "\n")
    return 0


if __name__ == "__main__":
    sys.exit(main())
