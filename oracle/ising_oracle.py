"""NumPy restatement of pyising's checkerboard Ising algorithm -- the correctness oracle.

TEST INFRASTRUCTURE ONLY: used by ``tests/``, the smoke check and the benchmark's CPU-baseline
leg as the checker; never imported by the product package.

Every function names the reference file:line it restates (paths relative to the reference
repository root). Pinning (what fails if a plausible mistake is made here):

* ``philox4x32`` -- the three Random123 known-answer vectors (tests/test_philox.py:7-12 of the
  reference) and the scalar-vs-vectorised check.
* ``split_seed`` / ``site_uniforms`` / ``word_randoms`` -- the stream-layout facts of
  tests/test_philox.py:29-34 and 45-49 of the reference, plus direct recomputation of single
  entries from ``philox4x32``.
* ``threshold_table`` -- closed-form values at beta = 0.5 (tests/test_lattice.py:33-38 of the
  reference) and the monotonicity / heat-bath symmetry t(h) + t(-h) = 2^32 - {0,1}.
* layouts -- round trips and the independent ``parity_grid`` selection
  (tests/test_lattice.py:19-30 of the reference).
* ``Int8Oracle`` -- ``naive_sweep``: a full-lattice restatement written from the definitions
  (tests/test_lattice.py:41-72 of the reference).
* ``MscOracle`` -- ``msc_eager_sweep``: full lattice, per-site uniform assembled from bit b of
  the 32 levels and compared with the int8 integer Metropolis rule (the eager formulation,
  src/pyising/backends/cuda/ising.cu:135 non-LAZY path), and the beta = 0 limit against int8.
* observables -- brute-force ``np.roll`` definitions and the cold-start ground state.
"""

from __future__ import annotations

import math

import numpy as np

# Philox4x32-10 constants: src/pyising/philox.py:12-17.
M0 = 0xD2511F53
M1 = 0xCD9E8D57
W0 = 0x9E3779B9
W1 = 0xBB67AE85
MASK32 = 0xFFFFFFFF
UINT32_MAX = 0xFFFFFFFF
ALL_ONES = np.uint64(0xFFFFFFFFFFFFFFFF)


# ------------------------------------------------------------------------------ Philox stream --
def split_seed(seed: int) -> tuple[int, int]:
    """64-bit seed -> (k0 = low 32 bits, k1 = high 32 bits). src/pyising/philox.py:20-23."""
    s = int(seed) % (1 << 64)
    return s & MASK32, s >> 32


def philox4x32(c0, c1, c2, c3, k0: int, k1: int, rounds: int = 10):
    """Vectorised Philox4x32-R on uint64-emulated 32-bit words. src/pyising/philox.py:26-51.

    Round: (hi0, lo0) = M0 * c0, (hi1, lo1) = M1 * c2;
    (c0, c1, c2, c3) <- (hi1 ^ c1 ^ k0, lo1, hi0 ^ c3 ^ k1, lo0); then the key is bumped by
    (W0, W1). Returns four uint32 arrays.
    """
    mask = np.uint64(MASK32)
    x = [np.asarray(c, dtype=np.uint64) & mask for c in (c0, c1, c2, c3)]
    x = list(np.broadcast_arrays(*x))
    ka, kb = int(k0) & MASK32, int(k1) & MASK32
    m0, m1, sh = np.uint64(M0), np.uint64(M1), np.uint64(32)
    for _ in range(rounds):
        p0 = x[0] * m0
        p1 = x[2] * m1
        x = [(p1 >> sh) ^ x[1] ^ np.uint64(ka), p1 & mask, (p0 >> sh) ^ x[3] ^ np.uint64(kb),
             p0 & mask]
        ka = (ka + W0) & MASK32
        kb = (kb + W1) & MASK32
    return tuple(v.astype(np.uint32) for v in x)


def site_uniforms(n_sites: int, t: int, color: int, seed: int) -> np.ndarray:
    """int8 stream: colour-site n gets word n % 4 of Philox((n // 4, t, color, 0), key).

    src/pyising/philox.py:54-64; the quad counter is truncated to 32 bits as in
    src/pyising/backends/cuda/ising.cu:78 (``(u32)q``).
    """
    if n_sites % 4:
        raise ValueError("number of colour sites must be a multiple of 4")
    k0, k1 = split_seed(seed)
    q = np.arange(n_sites // 4, dtype=np.uint64)
    r = philox4x32(q, t, color, 0, k0, k1)
    return np.stack(r, axis=1).reshape(-1)


def word_randoms(n_words: int, t: int, color: int, seed: int, level_pair: int):
    """msc stream: word w, levels 2l and 2l+1 from Philox((w, t, color, 1 + l), key).

    Level 2l is ``r0 | r1 << 32``, level 2l+1 is ``r2 | r3 << 32``. src/pyising/philox.py:67-78;
    the word counter is truncated to 32 bits as in src/pyising/backends/cuda/ising.cu:137.
    """
    k0, k1 = split_seed(seed)
    w = np.arange(n_words, dtype=np.uint64)
    r0, r1, r2, r3 = (v.astype(np.uint64) for v in philox4x32(w, t, color, 1 + level_pair, k0, k1))
    s = np.uint64(32)
    return r0 | (r1 << s), r2 | (r3 << s)


# ------------------------------------------------------------------------------- thresholds --
def prob_to_threshold(p: float) -> int:
    """floor(p * 2^32) clamped to UINT32_MAX; p >= 1 -> UINT32_MAX. src/pyising/lattice.py:121-125."""
    if p >= 1.0:
        return UINT32_MAX
    return min(int(math.floor(p * 2.0**32)), UINT32_MAX)


def threshold_table(beta: float, ndim: int, algorithm: str) -> np.ndarray:
    """uint32 table indexed by (X + 2 ndim) // 2 for X in -2d..2d step 2. src/pyising/lattice.py:128-142.

    Metropolis (X = s h): UINT32_MAX for X <= 0, else floor(exp(-2 beta X) 2^32).
    Heat bath (X = h): floor(2^32 / (1 + exp(-2 beta h))).
    """
    z = 2 * ndim
    xs = list(range(-z, z + 1, 2))
    if algorithm == "metropolis":
        vals = [UINT32_MAX if x <= 0 else prob_to_threshold(math.exp(-2.0 * beta * x)) for x in xs]
    elif algorithm == "heatbath":
        vals = [prob_to_threshold(1.0 / (1.0 + math.exp(-2.0 * beta * x))) for x in xs]
    else:
        raise ValueError(f"unknown algorithm {algorithm!r}")
    return np.array(vals, dtype=np.uint32)


# ---------------------------------------------------------------------------------- layouts --
def parity_grid(shape) -> np.ndarray:
    """(sum of indices) % 2 of every site. src/pyising/lattice.py:37-40."""
    return (np.indices(shape).sum(axis=0) % 2).astype(np.uint8)


def row_parity(shape) -> np.ndarray:
    """Parity of the non-x index sum of each flattened row. src/pyising/lattice.py:43-46."""
    return (np.indices(shape[:-1]).sum(axis=0) % 2).reshape(-1).astype(np.int64)


def to_compressed(lat: np.ndarray):
    """Full lattice -> (c0, c1), each (rows, Lx/2); colour c of row r sits at x = 2j + (p_r + c) % 2.

    src/pyising/lattice.py:49-60 (restated with strided slices instead of gathers).
    """
    lx = lat.shape[-1]
    flat = lat.reshape(-1, lx)
    odd = (row_parity(lat.shape) == 1)[:, None]
    even_cols, odd_cols = flat[:, 0::2], flat[:, 1::2]
    c0 = np.ascontiguousarray(np.where(odd, odd_cols, even_cols))
    c1 = np.ascontiguousarray(np.where(odd, even_cols, odd_cols))
    return c0, c1


def from_compressed(c0: np.ndarray, c1: np.ndarray, shape) -> np.ndarray:
    """Inverse of :func:`to_compressed`. src/pyising/lattice.py:63-72."""
    lx = shape[-1]
    odd = (row_parity(shape) == 1)[:, None]
    flat = np.empty((c0.shape[0], lx), dtype=c0.dtype)
    flat[:, 0::2] = np.where(odd, c1, c0)
    flat[:, 1::2] = np.where(odd, c0, c1)
    return flat.reshape(shape)


def pack_bits(comp: np.ndarray) -> np.ndarray:
    """(rows, Lh) int8 -> (rows, Lh/64) uint64, bit b of word w <-> j = 64 w + b, 1 = +1.

    src/pyising/lattice.py:75-82.
    """
    rows, lh = comp.shape
    if lh % 64:
        raise ValueError("Lx/2 must be a multiple of 64 for the packed layout")
    by = np.packbits(comp > 0, axis=1, bitorder="little")
    return np.ascontiguousarray(by).view("<u8").astype(np.uint64).reshape(rows, lh // 64)


def unpack_bits(words: np.ndarray) -> np.ndarray:
    """Inverse of :func:`pack_bits`, returns int8 +-1. src/pyising/lattice.py:85-90."""
    rows, nw = words.shape
    by = np.ascontiguousarray(words.astype("<u8")).view(np.uint8).reshape(rows, nw * 8)
    bits = np.unpackbits(by, axis=1, bitorder="little").astype(np.int8)
    return bits * 2 - 1


def random_lattice(shape, seed: int) -> np.ndarray:
    """Hot start: numpy default_rng(seed).integers(0, 2, shape, int8) * 2 - 1 (PCG64).

    src/pyising/lattice.py:93-96.
    """
    rng = np.random.default_rng(seed)
    a = rng.integers(0, 2, size=tuple(shape), dtype=np.int8)
    a *= 2
    a -= 1
    return a


# ----------------------------------------------------------------- full-lattice definitions --
def neighbor_sum(lat: np.ndarray) -> np.ndarray:
    """Sum of the 2d periodic nearest neighbours. src/pyising/lattice.py:99-105."""
    h = np.zeros(lat.shape, dtype=np.int64)
    for ax in range(lat.ndim):
        h += np.roll(lat, 1, axis=ax)
        h += np.roll(lat, -1, axis=ax)
    return h


def observables_exact(lat: np.ndarray) -> tuple[int, int]:
    """Exact integers (M, E): M = sum s, E = -sum over bonds s_i s_j. src/pyising/lattice.py:108-118."""
    m = int(np.sum(lat, dtype=np.int64))
    e = 0
    for ax in range(lat.ndim):
        e -= int(np.sum(lat.astype(np.int64) * np.roll(lat, 1, axis=ax)))
    return m, e


def naive_sweep(lat: np.ndarray, beta: float, seed: int, t: int, algorithm: str) -> np.ndarray:
    """One full sweep on the full lattice, written from the definitions.

    Restates tests/test_lattice.py:41-72 of the reference: colour c sites, in C order of the
    full lattice, are colour-site indices n = 0, 1, ...; u(n) from :func:`site_uniforms`.
    """
    lat = lat.copy()
    d = lat.ndim
    par = parity_grid(lat.shape)
    table = threshold_table(beta, d, algorithm)
    for color in (0, 1):
        h = neighbor_sum(lat)
        sel = par == color
        u = site_uniforms(lat.size // 2, t, color, seed)
        s = lat[sel].astype(np.int64)
        x = h[sel] if algorithm == "heatbath" else s * h[sel]
        acc = u < table[(x + 2 * d) // 2]
        if algorithm == "heatbath":
            new = np.where(acc, 1, -1)
        else:
            new = np.where(acc | (x <= 0), -s, s)
        lat[sel] = new.astype(np.int8)
    return lat


def msc_site_uniform(shape, t: int, color: int, seed: int) -> np.ndarray:
    """Full 32-bit uniform of every colour-site in the msc stream (eager formulation).

    Bit (31 - k) of site j's uniform is bit (j % 64) of level k of word j // 64
    (src/pyising/backends/numpy_backend.py:73-82 with every level drawn), returned in
    colour-site order n = r * Lh + j.
    """
    rows = int(np.prod(shape[:-1]))
    lh = shape[-1] // 2
    nw = lh // 64
    n_words = rows * nw
    u = np.zeros((n_words, 64), dtype=np.uint64)
    lanes = np.arange(64, dtype=np.uint64)
    for lp in range(16):
        lv = word_randoms(n_words, t, color, seed, lp)
        for i in range(2):
            k = 2 * lp + i
            bits = (lv[i][:, None] >> lanes[None, :]) & np.uint64(1)
            u |= bits << np.uint64(31 - k)
    return u.astype(np.uint32).reshape(-1)


def msc_eager_sweep(lat: np.ndarray, beta: float, seed: int, t: int) -> np.ndarray:
    """One full msc sweep from the definitions: int8 Metropolis with the msc-stream uniform.

    The lazy bitwise comparison (src/pyising/backends/numpy_backend.py:54-83) decides exactly
    ``u < table[idx]`` with u assembled MSB-first from the level bits, so the eager form is
    an independent statement of the same rule (2D only).
    """
    lat = lat.copy()
    par = parity_grid(lat.shape)
    table = threshold_table(beta, 2, "metropolis")
    for color in (0, 1):
        h = neighbor_sum(lat)
        sel = par == color
        u = msc_site_uniform(lat.shape, t, color, seed)
        s = lat[sel].astype(np.int64)
        x = s * h[sel]
        flip = (x <= 0) | (u < table[(x + 4) // 2])
        lat[sel] = np.where(flip, -s, s).astype(np.int8)
    return lat


# --------------------------------------------------------------------------- compressed oracles --
def check_shape(shape, multiple: int = 8):
    """2D/3D, even extents >= 2, last extent a multiple of ``multiple``. src/pyising/lattice.py:25-34."""
    shape = tuple(int(s) for s in shape)
    if len(shape) not in (2, 3):
        raise ValueError(f"only 2D and 3D lattices are supported, got shape {shape}")
    if any(s < 2 or s % 2 for s in shape):
        raise ValueError(f"all extents must be even and >= 2, got {shape}")
    if shape[-1] % multiple:
        raise ValueError(f"last extent must be a multiple of {multiple}, got {shape[-1]}")
    return shape


class Int8Oracle:
    """Compressed-checkerboard int8 simulation. src/pyising/backends/numpy_backend.py:14-51
    with the sweep loop of src/pyising/backends/base.py:85-90."""

    def __init__(self, shape, beta, seed=0, algorithm="metropolis", init="hot"):
        self.shape = check_shape(shape, 8)
        self.ndim = len(self.shape)
        self.algorithm = algorithm
        self.seed = int(seed) % (1 << 64)
        self.beta = float(beta)
        self.t = 0
        self.rows = int(np.prod(self.shape[:-1]))
        self.lh = self.shape[-1] // 2
        self.table = threshold_table(self.beta, self.ndim, algorithm)
        self.odd = (row_parity(self.shape) == 1)[:, None]
        if isinstance(init, np.ndarray):
            self.set_lattice(init)
        elif init == "cold":
            self.set_lattice(np.ones(self.shape, np.int8))
        elif init == "hot":
            self.set_lattice(random_lattice(self.shape, self.seed))
        else:
            raise ValueError(f"unknown init {init!r}")

    def set_beta(self, beta):
        self.beta = float(beta)
        self.table = threshold_table(self.beta, self.ndim, self.algorithm)

    def set_lattice(self, lat):
        self.c = list(to_compressed(np.asarray(lat, np.int8).reshape(self.shape)))

    def get_lattice(self):
        return from_compressed(self.c[0], self.c[1], self.shape)

    def neighbor_sum(self, color):
        """h of every colour-``color`` site. src/pyising/backends/numpy_backend.py:26-33."""
        o = self.c[1 - color].astype(np.int64)
        nxt = np.roll(o, -1, axis=1)  # j + 1
        prv = np.roll(o, 1, axis=1)   # j - 1
        p1 = self.odd if color == 0 else ~self.odd  # p = (rowparity + color) % 2 == 1
        h = o + np.where(p1, nxt, prv)
        g = o.reshape(self.shape[:-1] + (self.lh,))
        for ax in range(self.ndim - 1):
            h += (np.roll(g, 1, axis=ax) + np.roll(g, -1, axis=ax)).reshape(self.rows, self.lh)
        return h

    def half_sweep(self, color):
        """src/pyising/backends/numpy_backend.py:35-44."""
        s = self.c[color].astype(np.int64)
        h = self.neighbor_sum(color)
        u = site_uniforms(self.rows * self.lh, self.t, color, self.seed).reshape(s.shape)
        x = h if self.algorithm == "heatbath" else s * h
        acc = u < self.table[(x + 2 * self.ndim) // 2]
        if self.algorithm == "heatbath":
            self.c[color] = np.where(acc, 1, -1).astype(np.int8)
        else:
            self.c[color] = np.where(acc | (x <= 0), -s, s).astype(np.int8)

    def sweep(self, n=1):
        for _ in range(int(n)):
            self.half_sweep(0)
            self.half_sweep(1)
            self.t += 1

    def observables_exact(self):
        """(M, E) as exact integers; E from colour-0 bonds (src/pyising/backends/observables_torch.py:46-69)."""
        m = int(self.c[0].sum(dtype=np.int64) + self.c[1].sum(dtype=np.int64))
        e = -int((self.c[0].astype(np.int64) * self.neighbor_sum(0)).sum())
        return m, e

    def observables(self):
        m, e = self.observables_exact()
        n = int(np.prod(self.shape))
        return m / n, e / n


def msc_flip_mask(own, nbs, thr4: int, thr8: int, rand_level):
    """Bit-sliced Metropolis with the lazy MSB-first comparison.

    src/pyising/backends/numpy_backend.py:54-83: x_k = own ^ nb_k; a0 = no anti-aligned
    neighbour, a1 = exactly one, a2 = two or more; a2 lanes flip; a1 (a0) lanes flip iff
    u < thr4 (thr8) decided bit by bit.
    """
    x1, x2, x3, x4 = (own ^ nb for nb in nbs)
    anyx = x1 | x2 | x3 | x4
    a0 = ~anyx
    s1, c1 = x1 ^ x2, x1 & x2
    s2, c2 = x3 ^ x4, x3 & x4
    a2 = c1 | c2 | (s1 & s2)
    a1 = anyx & ~a2
    acc = np.zeros_like(own)
    und = a0 | a1
    for k in range(32):
        if not und.any():
            break
        r = rand_level(k)
        bit = 31 - k
        t4 = ALL_ONES if (thr4 >> bit) & 1 else np.uint64(0)
        t8 = ALL_ONES if (thr8 >> bit) & 1 else np.uint64(0)
        tm = (a1 & t4) | (a0 & t8)
        acc |= und & ~r & tm
        und &= ~(r ^ tm)
    return a2 | acc


class MscOracle:
    """Multi-spin-coded 2D Metropolis. src/pyising/backends/numpy_backend.py:86-131."""

    def __init__(self, shape, beta, seed=0, algorithm="metropolis", init="hot"):
        self.shape = check_shape(shape, 128)
        if len(self.shape) != 2:
            raise ValueError("msc supports dims (2,)")
        if algorithm != "metropolis":
            raise ValueError("msc supports ('metropolis',)")
        self.ndim = 2
        self.algorithm = algorithm
        self.seed = int(seed) % (1 << 64)
        self.beta = float(beta)
        self.t = 0
        self.rows = self.shape[0]
        self.lh = self.shape[1] // 2
        self.nw = self.lh // 64
        self.table = threshold_table(self.beta, 2, "metropolis")
        self.odd = (row_parity(self.shape) == 1)[:, None]
        if isinstance(init, np.ndarray):
            self.set_lattice(init)
        elif init == "cold":
            self.w = [np.full((self.rows, self.nw), ALL_ONES, np.uint64) for _ in range(2)]
        elif init == "hot":
            self.set_lattice(random_lattice(self.shape, self.seed))
        else:
            raise ValueError(f"unknown init {init!r}")

    def set_beta(self, beta):
        self.beta = float(beta)
        self.table = threshold_table(self.beta, 2, "metropolis")

    def set_lattice(self, lat):
        c0, c1 = to_compressed(np.asarray(lat, np.int8).reshape(self.shape))
        self.w = [pack_bits(c0), pack_bits(c1)]

    def get_lattice(self):
        return from_compressed(unpack_bits(self.w[0]), unpack_bits(self.w[1]), self.shape)

    def neighbors(self, color):
        """[same j, horizontal, y-1, y+1] words. src/pyising/backends/numpy_backend.py:103-110."""
        o = self.w[1 - color]
        one, s63 = np.uint64(1), np.uint64(63)
        right = (o >> one) | (np.roll(o, -1, axis=1) << s63)
        left = (o << one) | (np.roll(o, 1, axis=1) >> s63)
        p1 = self.odd if color == 0 else ~self.odd
        return [o, np.where(p1, right, left), np.roll(o, 1, axis=0), np.roll(o, -1, axis=0)]

    def half_sweep(self, color):
        """src/pyising/backends/numpy_backend.py:112-124."""
        own = self.w[color]
        n_words = self.rows * self.nw
        cache = {}

        def rand_level(k):
            if k // 2 not in cache:
                cache[k // 2] = word_randoms(n_words, self.t, color, self.seed, k // 2)
            return cache[k // 2][k % 2].reshape(own.shape)

        flip = msc_flip_mask(own, self.neighbors(color), int(self.table[3]), int(self.table[4]),
                             rand_level)
        self.w[color] = own ^ flip

    def sweep(self, n=1):
        for _ in range(int(n)):
            self.half_sweep(0)
            self.half_sweep(1)
            self.t += 1

    def observables_exact(self):
        """(M, E) exact: M = 2 * ones - N, E = -(2 N - 2 anti) from colour-0 words.

        src/pyising/backends/observables_torch.py:85-111.
        """
        n = self.rows * self.nw * 128
        ones = int(sum(int(np.unpackbits(w.view(np.uint8)).sum()) for w in self.w))
        s = self.w[0]
        anti = 0
        for nb in self.neighbors(0):
            anti += int(np.unpackbits((s ^ nb).view(np.uint8)).sum())
        return 2 * ones - n, -(2 * n - 2 * anti)

    def observables(self):
        m, e = self.observables_exact()
        n = self.rows * self.nw * 128
        return m / n, e / n
