"""Host-side lattice conventions of the hip backend: shape rules, acceptance tables, seeds,
hot-start draws, and exact reference constants.

These are the host halves of the drop-in boundary. They follow the reference's conventions
(src/pyising/lattice.py, src/pyising/philox.py, src/pyising/exact.py) so that the same
(shape, beta, seed) produce the same integer thresholds, Philox keys and hot-start spins; the
conversions between full lattices and colour arrays run on the GPU
(csrc/layout_kernels.hip).
"""

from __future__ import annotations

import math

import numpy as np

UINT32_MAX = 0xFFFFFFFF

#: Onsager: sinh(2 / T_c) = 1 (src/pyising/exact.py:10-12)
TC_2D = 2.0 / math.log(1.0 + math.sqrt(2.0))
BETA_C_2D = 1.0 / TC_2D
#: simple-cubic critical coupling, Ferrenberg, Xu & Landau 2018 (src/pyising/exact.py:13-16)
BETA_C_3D = 0.221654626
TC_3D = 1.0 / BETA_C_3D


def check_shape(shape, multiple: int = 8) -> tuple[int, ...]:
    """2D or 3D, every extent even and >= 2, last extent a multiple of ``multiple``.

    Raises ``ValueError`` like src/pyising/lattice.py:25-34.
    """
    shape = tuple(int(s) for s in shape)
    if len(shape) not in (2, 3):
        raise ValueError(f"only 2D and 3D lattices are supported, got shape {shape}")
    if any(s < 2 or s % 2 for s in shape):
        raise ValueError(f"all extents must be even and >= 2, got {shape}")
    if shape[-1] % multiple:
        raise ValueError(f"last extent must be a multiple of {multiple}, got {shape[-1]}")
    return shape


def split_seed(seed: int) -> tuple[int, int]:
    """64-bit seed -> Philox key (lo32, hi32) (src/pyising/philox.py:20-23)."""
    s = int(seed) & 0xFFFFFFFFFFFFFFFF
    return s & 0xFFFFFFFF, s >> 32


def prob_to_threshold(p: float) -> int:
    """Accept iff u32 < threshold; floor(p 2^32), p >= 1 -> UINT32_MAX (src/pyising/lattice.py:121-125)."""
    if p >= 1.0:
        return UINT32_MAX
    return min(int(math.floor(p * 2.0**32)), UINT32_MAX)


def threshold_table(beta: float, ndim: int, algorithm: str) -> np.ndarray:
    """uint32 table of length 2 ndim + 1, indexed by (X + 2 ndim) // 2 (src/pyising/lattice.py:128-142).

    Computed on the host in float64 with ``math.exp`` so the integers match the reference
    exactly; the kernels only ever compare against these integers.
    """
    z = 2 * ndim
    xs = range(-z, z + 1, 2)
    if algorithm == "metropolis":
        vals = [UINT32_MAX if x <= 0 else prob_to_threshold(math.exp(-2.0 * beta * x)) for x in xs]
    elif algorithm == "heatbath":
        vals = [prob_to_threshold(1.0 / (1.0 + math.exp(-2.0 * beta * x))) for x in xs]
    else:
        raise ValueError(f"unknown algorithm {algorithm!r}")
    return np.array(vals, dtype=np.uint32)


def pcg64_state(seed: int) -> tuple[int, int]:
    """The seeded (state, inc) of numpy's PCG64 behind ``default_rng(seed)`` (numpy's own
    SeedSequence seeding); the device hot-start kernel continues this stream."""
    st = np.random.PCG64(int(seed) & 0xFFFFFFFFFFFFFFFF).state["state"]
    return int(st["state"]), int(st["inc"])


def hot_start_bits(shape, seed: int) -> np.ndarray:
    """The hot start's random draws: ``numpy.random.default_rng(seed).integers(0, 2, shape, int8)``.

    The reference's hot start is these draws mapped 0 -> -1, 1 -> +1
    (src/pyising/lattice.py:93-96); the device split kernels apply that map, so the {0, 1}
    array is uploaded as is.
    """
    rng = np.random.default_rng(int(seed) & 0xFFFFFFFFFFFFFFFF)
    return rng.integers(0, 2, size=tuple(shape), dtype=np.int8)


def row_parity(shape) -> np.ndarray:
    """Parity of the non-x index sum of each flattened row (src/pyising/lattice.py:43-46)."""
    return (np.indices(tuple(shape[:-1])).sum(axis=0) % 2).reshape(-1).astype(np.int64)
