"""Exact results for the square-lattice Ising model (J = k_B = 1), used as physics checks.

Same functions and meanings as the reference's src/pyising/exact.py:10-88, written from the
published closed forms:

* Onsager (1944): T_c = 2 / ln(1 + sqrt 2); internal energy per spin
  u = -coth 2K [1 + (2/pi)(2 tanh^2 2K - 1) K(k^2)], k = 2 sinh 2K / cosh^2 2K.
* Yang (1952): spontaneous magnetisation (1 - sinh(2K)^-4)^(1/8) below T_c.
* Kaufman (1949), Ferdinand & Fisher (1969): partition function of the m x n torus,
  Z = 1/2 (2 sinh 2K)^(mn/2) sum_{i=1..4} Z_i with
  Z_1,2 = prod_{r<n} 2 cosh / 2 sinh (m g_{2r+1} / 2), Z_3,4 = the same over g_{2r},
  cosh g_l = cosh 2K coth 2K - cos(pi l / n), and g_0 = 2K + ln tanh K (signed).
Everything is evaluated in log space so large tori do not overflow.
"""

from __future__ import annotations

import math

import numpy as np

TC_2D = 2.0 / math.log(1.0 + math.sqrt(2.0))
BETA_C_2D = 1.0 / TC_2D
BETA_C_3D = 0.221654626
BETA_C_3D_ERR = 0.000000005
TC_3D = 1.0 / BETA_C_3D


def magnetization_2d(T):
    """Spontaneous magnetisation per spin of the infinite lattice (0 above T_c)."""
    T = np.asarray(T, dtype=float)
    with np.errstate(divide="ignore", invalid="ignore"):
        s = np.sinh(2.0 / T)
        m = np.where(T < TC_2D, np.clip(1.0 - s**-4.0, 0.0, None) ** 0.125, 0.0)
    return m


def energy_2d(T):
    """Onsager's internal energy per spin of the infinite square lattice."""
    from scipy.special import ellipk

    K = 1.0 / np.asarray(T, dtype=float)
    k = 2.0 * np.sinh(2.0 * K) / np.cosh(2.0 * K) ** 2
    m = np.minimum(k * k, 1.0 - 1e-16)
    t2 = np.tanh(2.0 * K) ** 2
    return -(1.0 / np.tanh(2.0 * K)) * (1.0 + (2.0 / np.pi) * (2.0 * t2 - 1.0) * ellipk(m))


def _log_cosh2(x: np.ndarray) -> np.ndarray:
    """ln(2 cosh x)."""
    a = np.abs(x)
    return a + np.log1p(np.exp(-2.0 * a))


def _log_sinh2(x: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
    """(ln |2 sinh x|, sign(x))."""
    a = np.abs(x)
    with np.errstate(divide="ignore"):
        return a + np.log1p(-np.exp(-2.0 * a)), np.sign(x)


def log_partition_torus(K: float, m: int, n: int) -> float:
    """ln Z of the m x n periodic square lattice at coupling K = 1/T."""
    c = math.cosh(2.0 * K) / math.tanh(2.0 * K)

    def gam(l: np.ndarray) -> np.ndarray:
        g = np.arccosh(np.maximum(c - np.cos(np.pi * l / n), 1.0))
        return np.where(l == 0, 2.0 * K + math.log(math.tanh(K)), g)

    r = np.arange(n)
    parts = []  # (log |Z_i|, sign)
    for g in (gam(2 * r + 1), gam(2 * r)):
        parts.append((float(np.sum(_log_cosh2(0.5 * m * g))), 1.0))
        ls, sg = _log_sinh2(0.5 * m * g)
        parts.append((float(np.sum(ls)), float(np.prod(sg))))
    top = max(v for v, _ in parts if np.isfinite(v))
    total = sum(s * math.exp(v - top) for v, s in parts if np.isfinite(v))
    return -math.log(2.0) + 0.5 * m * n * math.log(2.0 * math.sinh(2.0 * K)) + top + math.log(total)


def energy_torus(T: float, L: int, h: float = 1e-5) -> float:
    """Exact energy per spin of the L x L torus, -d ln Z / dK / L^2 (central difference)."""
    K = 1.0 / T
    d = (log_partition_torus(K + h, L, L) - log_partition_torus(K - h, L, L)) / (2.0 * h)
    return -d / (L * L)


def specific_heat_torus(T: float, L: int, h: float = 1e-4) -> float:
    """Exact specific heat per spin of the L x L torus, K^2 d^2 ln Z / dK^2 / L^2."""
    K = 1.0 / T
    f = [log_partition_torus(K + s * h, L, L) for s in (-1, 0, 1)]
    return K * K * (f[0] - 2.0 * f[1] + f[2]) / (h * h) / (L * L)
