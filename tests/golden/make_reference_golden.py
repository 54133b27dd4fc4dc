"""Writes ``reference_msc_vectors.json``: vectors produced by the reference's own Python code for
the multi-spin-coded (msc) random stream and update, the one part of the path that no known
answer of the reference's tests and no closed form pins.

    python tests/golden/make_reference_golden.py      (in the build container only)

It imports, from the reference checkout at /root/reference/src, exactly two things:

* ``pyising.philox.word_randoms`` (src/pyising/philox.py:67-78): the two uint64 level words
  (levels 2l and 2l + 1) per packed word, for chosen (sweep, colour, seed, level pair);
* ``pyising.backends.numpy_backend.NumpyMultiSpin`` (src/pyising/backends/numpy_backend.py:
  86-131) with its start lattices from ``pyising.lattice.random_lattice``
  (src/pyising/lattice.py:93-96): end states (sha256 of the full int8 lattice in C order, exact
  M and E) after the sweeps of the reference's own msc test cases (tests/test_backends.py:
  40-68, src/pyising/validate.py:72-79) and of the configured msc workload's shape.

The reference never leaves this container: only the JSON data travels with the tests, which
check the oracle (CPU) and the hip backend (GPU) against it. numpy's version is recorded
because the hot starts come from numpy's default_rng.
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
REFERENCE_SRC = Path("/root/reference/src")
OUT = HERE / "reference_msc_vectors.json"

BETA_C_2D = 0.5 * math.log(1.0 + math.sqrt(2.0))

#: (sweep, colour, seed, level pair) of the stream vectors; seeds past 2^32 exercise the key's
#: high word, sweeps past 2^32 the 32-bit counter truncation, level pairs 0 and 15 both ends
STREAM_CASES = [
    (0, 0, 0, 0),
    (0, 1, 1, 0),
    (5, 1, 2**40 + 3, 0),
    (5, 1, 2**40 + 3, 15),
    (123456, 0, 2**63 + 12345, 7),
    (2**32 - 1, 1, 1, 3),
    (2**32 + 5, 0, 11, 1),
    (7, 1, 0xFFFFFFFFFFFFFFFF, 15),
    (999, 0, 99, 8),
]
STREAM_WORDS = 70  # word indices 0..69: past one 64-lane wave


def _digest(lat: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(lat, dtype=np.int8).tobytes()).hexdigest()


def _exact_m_e(lat: np.ndarray) -> tuple[int, int]:
    """M = sum of spins, E = -(sum over bonds of s_i s_j), from the definition
    (periodic, every bond once)."""
    s = lat.astype(np.int64)
    bonds = sum(int((s * np.roll(s, -1, axis=ax)).sum()) for ax in range(s.ndim))
    return int(s.sum()), -bonds


def stream_vectors(word_randoms) -> list[dict]:
    out = []
    for sweep, color, seed, lp in STREAM_CASES:
        lo, hi = word_randoms(STREAM_WORDS, sweep, color, seed, lp)
        out.append({"sweep": sweep, "color": color, "seed": seed, "level_pair": lp,
                    "level_even": [f"{int(x):016x}" for x in lo],
                    "level_odd": [f"{int(x):016x}" for x in hi]})
    return out


def end_state(cls, random_lattice, name, shape, beta, seed, sweeps, init, init_seed=None,
              then=None) -> dict:
    t0 = time.time()
    shape = tuple(shape)
    start = random_lattice(shape, init_seed) if init == "random" else init
    sim = cls(shape, beta, seed=seed, init=start)
    table = [int(x) for x in sim.table]
    sim.sweep(sweeps)
    if then is not None:
        sim.set_beta(then[0])
        sim.sweep(then[1])
    lat = sim.get_lattice()
    M, E = _exact_m_e(lat)
    print(f"{name}: M={M} E={E} ({time.time() - t0:.1f} s)", flush=True)
    return {"name": name, "shape": list(shape), "beta": beta, "seed": seed, "sweeps": sweeps,
            "init": init, "init_seed": init_seed, "then": list(then) if then else None,
            "thr4": table[3], "thr8": table[4], "sha256": _digest(lat), "M": M, "E": E}


def end_states(cls, random_lattice) -> list[dict]:
    S = []

    def add(*a, **k):
        S.append(end_state(cls, random_lattice, *a, **k))

    # the reference's msc identity cases
    add("identity_msc_128x128", (128, 128), 0.44, 11, 5, "random", 9)      # test_backends.py:40-52
    add("identity_msc_64x384", (64, 384), 0.44, 11, 5, "random", 9)
    add("set_beta_msc_128x128", (128, 128), 0.3, 3, 2, "random", 1, then=(0.5, 3))  # :55-68
    add("validate_msc_128x256", (128, 256), 0.42, 99, 8, "random", 2024)   # validate.py:72-79
    # odd word counts per row (nw = 1, 3, 5) and temperatures whose thresholds differ in bits
    add("msc_64x128_beta0.6", (64, 128), 0.6, 5, 12, "random", 3)
    add("msc_130x384_beta0.3", (130, 384), 0.3, 2**40 + 9, 10, "random", 4)
    add("msc_96x640_beta_c", (96, 640), BETA_C_2D, 2**33 + 1, 15, "random", 6)
    # the configured workload's shape family: hot start, T_c, seed 1
    add("msc_128x256_beta_c_hot_20", (128, 256), BETA_C_2D, 1, 20, "hot")
    add("msc_L1024_beta_c_hot_200", (1024, 1024), BETA_C_2D, 1, 200, "hot")
    add("msc_L16384_beta_c_hot_3", (16384, 16384), BETA_C_2D, 1, 3, "hot")
    return S


def main() -> int:
    sys.dont_write_bytecode = True  # the reference checkout stays untouched (no __pycache__)
    sys.path.insert(0, str(REFERENCE_SRC))
    from pyising.backends.numpy_backend import NumpyMultiSpin
    from pyising.lattice import random_lattice
    from pyising.philox import word_randoms

    data = {
        "source": "the reference's own NumPy code (src/pyising/philox.py:67-78 word_randoms; "
                  "src/pyising/backends/numpy_backend.py:86-131 NumpyMultiSpin; "
                  "src/pyising/lattice.py:93-96 random_lattice), run in the build container",
        "numpy": np.__version__,
        "note": "level words as 16-hex-digit uint64; sha256 over the full int8 lattice (+1/-1) "
                "in C order; M, E exact integers from the definition",
        "stream_words": STREAM_WORDS,
        "stream": stream_vectors(word_randoms),
        "end_states": end_states(NumpyMultiSpin, random_lattice),
    }
    OUT.write_text(json.dumps(data, indent=1) + "\n")
    print(f"wrote {OUT}")
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
