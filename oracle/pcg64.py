"""TEST INFRASTRUCTURE ONLY (the oracle): a pure-Python restatement of the random stream behind
the reference's hot start, used by tests to pin the device hot-start generator.

The reference's hot start is ``numpy.random.default_rng(seed).integers(0, 2, shape, int8)``
(src/pyising/lattice.py:93-96). That stream lives in numpy (third-party, numpy 2.2 here;
unpinned ``>=1.24`` in the reference's pyproject.toml:14), restated from its published algorithm:

* ``default_rng(seed)`` is ``Generator(PCG64(seed))``; the seeding (SeedSequence) is taken from
  numpy itself (``PCG64(seed).state``), not restated.
* PCG64 is the PCG XSL-RR 128/64 generator (O'Neill 2014): state <- state * M + inc (mod 2^128)
  with M = 0x2360ED051FC65DA44385DF649FCCF645, output = rotr64(hi ^ lo, state >> 122) of the new
  state.
* ``integers(0, 2, dtype=int8)`` draws buffered bytes from 32-bit words, the low half of each
  64-bit output first, and Lemire's bounded method for range 2 keeps byte >> 7 (no rejection:
  its threshold is (255 - 1) % 2 = 0). Byte i is therefore bit 8 (i % 8) + 7 of 64-bit output
  i // 8.
* jump-ahead: the LCG advance by repeated squaring (Brown 1994).

Pinned in tests/test_oracle_pcg64.py against numpy's own ``random_raw``, ``advance`` and
``integers``.
"""

from __future__ import annotations

import numpy as np

MULT = 0x2360ED051FC65DA44385DF649FCCF645
MASK128 = (1 << 128) - 1
MASK64 = (1 << 64) - 1


def seed_state(seed: int) -> tuple[int, int]:
    """(state, inc) of numpy's PCG64 for ``default_rng(seed)`` (numpy's own seeding)."""
    st = np.random.PCG64(int(seed) & MASK64).state["state"]
    return int(st["state"]), int(st["inc"])


def step(state: int, inc: int) -> int:
    return (state * MULT + inc) & MASK128


def output(state: int) -> int:
    hi, lo = state >> 64, state & MASK64
    x = hi ^ lo
    r = state >> 122
    return ((x >> r) | (x << (64 - r))) & MASK64


def advance(state: int, inc: int, delta: int) -> int:
    """The state after ``delta`` steps, in O(log delta)."""
    cur_mult, cur_plus = MULT, inc
    acc_mult, acc_plus = 1, 0
    while delta:
        if delta & 1:
            acc_mult = (acc_mult * cur_mult) & MASK128
            acc_plus = (acc_plus * cur_mult + cur_plus) & MASK128
        cur_plus = ((cur_mult + 1) * cur_plus) & MASK128
        cur_mult = (cur_mult * cur_mult) & MASK128
        delta >>= 1
    return (acc_mult * state + acc_plus) & MASK128


def raw(seed: int, n: int, start: int = 0) -> list[int]:
    """64-bit outputs start .. start + n - 1 of default_rng(seed)'s bit generator."""
    s, inc = seed_state(seed)
    s = advance(s, inc, start)
    out = []
    for _ in range(n):
        s = step(s, inc)
        out.append(output(s))
    return out


def hot_start_bytes(seed: int, n: int) -> np.ndarray:
    """``default_rng(seed).integers(0, 2, n, int8)`` from the restated stream."""
    words = raw(seed, (n + 7) // 8)
    out = np.empty(n, dtype=np.int8)
    for i in range(n):
        out[i] = (words[i // 8] >> (8 * (i % 8) + 7)) & 1
    return out
