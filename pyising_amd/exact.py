"""Exact results for the square-lattice Ising model (J = k_B = 1), the secondary physics checks.

The public names follow the reference's src/pyising/exact.py (TC_2D, BETA_C_2D, BETA_C_3D,
TC_3D, magnetization_2d, energy_2d, log_partition_torus, energy_torus, specific_heat_torus), so
callers of that module get the same quantities. The evaluation here is this module's own:

* Onsager's internal energy per spin of the infinite lattice,
  u(K) = -coth 2K [1 + (2/pi) (2 tanh^2 2K - 1) K(kappa)], kappa = 2 sinh 2K / cosh^2 2K,
  with the complete elliptic integral K(kappa) = pi / (2 AGM(1, sqrt(1 - kappa^2))) computed by
  the arithmetic-geometric mean (no special-function library); at T_c the factor
  2 tanh^2 2K - 1 vanishes where K(kappa) diverges, and the product is taken as its limit, 0.
* Yang's spontaneous magnetisation (1 - sinh(2K)^-4)^(1/8) below T_c.
* The m x n torus (Kaufman 1949): Z = 1/2 (2 sinh 2K)^(mn/2) (P1 + P2 + P3 + P4) with the four
  products P = prod_{r<n} F(m g_l / 2) over l = 2r+1 (P1: F = 2 cosh, P2: F = 2 sinh) and
  l = 2r (P3: 2 cosh, P4: 2 sinh), cosh g_l = cosh 2K coth 2K - cos(pi l / n) for l > 0 and the
  signed g_0 = 2K + ln tanh K. The energy and the specific heat are the analytic first and
  second K-derivatives of ln Z (no finite differences): for a product of F(x_r) with
  x_r = m g_r / 2, P'/P = sum_r a_r with a_r = (m/2) g_r' F'/F, and, since F''/F = 1 for both
  cosh and sinh, P''/P = sum_r ((m/2)^2 g_r'^2 + (m/2) g_r'' F'/F) + 2 sum_{r<s} a_r a_s. The
  derivatives of g follow from cosh g = c(K) - cos phi: g' = c' / sinh g,
  g'' = (c'' - cosh g g'^2) / sinh g, with c = cosh^2 2K / sinh 2K; for the signed g_0,
  g_0' = 2 + 2 / sinh 2K and g_0'' = -4 cosh 2K / sinh^2 2K. The four products are combined
  in a scaled form (each as its sign times exp(ln|P| - max)); in a product with a factor
  2 sinh(x) near x = 0 (g_0 at T_c) that factor is kept apart and enters by the product rule,
  so nothing is divided by it.
"""

from __future__ import annotations

import math

import numpy as np

#: Onsager (1944): sinh(2 / T_c) = 1.
TC_2D = 2.0 / math.log(1.0 + math.sqrt(2.0))
BETA_C_2D = 1.0 / TC_2D
#: Simple-cubic critical coupling (Ferrenberg, Xu & Landau, Phys. Rev. E 97, 043301, 2018).
BETA_C_3D = 0.221654626
BETA_C_3D_ERR = 0.000000005
TC_3D = 1.0 / BETA_C_3D


# ------------------------------------------------------------------ infinite lattice ----
def magnetization_2d(T):
    """Yang's spontaneous magnetisation per spin of the infinite lattice (0 at and above T_c)."""
    T = np.asarray(T, dtype=float)
    with np.errstate(divide="ignore", invalid="ignore", over="ignore"):
        inv4 = np.sinh(2.0 / T) ** -4.0
        m = np.where(T < TC_2D, np.maximum(1.0 - inv4, 0.0) ** 0.125, 0.0)
    return m


def _agm(a, b):
    """Arithmetic-geometric mean, elementwise, to double precision."""
    a, b = np.asarray(a, dtype=float).copy(), np.asarray(b, dtype=float).copy()
    for _ in range(64):
        a, b = 0.5 * (a + b), np.sqrt(a * b)
        if np.all(np.abs(a - b) <= 1e-16 * np.abs(a)):
            break
    return a


def elliptic_k(kappa):
    """Complete elliptic integral of the first kind K(kappa) (modulus kappa, 0 <= kappa < 1)."""
    kappa = np.asarray(kappa, dtype=float)
    return 0.5 * np.pi / _agm(np.ones_like(kappa), np.sqrt(1.0 - kappa * kappa))


def energy_2d(T):
    """Onsager's internal energy per spin of the infinite square lattice."""
    K = 1.0 / np.asarray(T, dtype=float)
    s2, c2 = np.sinh(2.0 * K), np.cosh(2.0 * K)
    kappa = np.minimum(2.0 * s2 / (c2 * c2), 1.0)
    lead = 2.0 * np.tanh(2.0 * K) ** 2 - 1.0
    near_tc = 1.0 - kappa * kappa < 1e-24  # K(kappa) infinite; lead vanishes linearly there
    with np.errstate(divide="ignore", invalid="ignore"):
        ek = elliptic_k(np.where(near_tc, 0.0, kappa))
    term = np.where(near_tc, 0.0, lead * ek)
    return -(c2 / s2) * (1.0 + (2.0 / np.pi) * term)


# ---------------------------------------------------------------------- the m x n torus ----
def _gamma(K: float, n: int, odd: bool):
    """g_l, g_l', g_l'' for l = 2r + 1 (odd) or l = 2r (even), r = 0 .. n-1."""
    s, ch = math.sinh(2.0 * K), math.cosh(2.0 * K)
    c = ch * ch / s
    c1 = 2.0 * ch * (s * s - 1.0) / (s * s)                 # dc/dK
    c2 = 4.0 * s * (1.0 - 1.0 / (s * s)) + 8.0 * ch * ch / s**3  # d2c/dK2
    l = 2 * np.arange(n) + (1 if odd else 0)
    cosh_g = c - np.cos(np.pi * l / n)
    g = np.arccosh(np.maximum(cosh_g, 1.0))
    sinh_g = np.sqrt(np.maximum(cosh_g * cosh_g - 1.0, 0.0))
    with np.errstate(divide="ignore", invalid="ignore"):
        g1 = c1 / sinh_g
        g2 = (c2 - cosh_g * g1 * g1) / sinh_g
    if not odd:  # l = 0: the signed branch g_0 = 2K + ln tanh K
        g[0] = 2.0 * K + math.log(math.tanh(K))
        g1[0] = 2.0 + 2.0 / s
        g2[0] = -4.0 * ch / (s * s)
    return g, g1, g2


def _product(m: int, g, g1, g2, sinh_type: bool):
    """P = prod_r F(m g_r / 2) (F = 2 cosh, or 2 sinh if sinh_type) and its first two
    K-derivatives as (log_scale, sign, v0, v1, v2) with P = sign exp(log_scale) v0,
    P' = sign exp(log_scale) v1, P'' = sign exp(log_scale) v2. Normally v0 = 1 and v1, v2 are
    P'/P, P''/P; a 2 sinh factor near x = 0 stays out of the scale and its value and
    derivatives enter v0, v1, v2 directly (product rule), so nothing is divided by it."""
    x = 0.5 * m * g
    h1, h2 = 0.5 * m * g1, 0.5 * m * g2
    ax = np.abs(x)
    keep = np.ones(len(x), bool)
    lin = None
    if sinh_type:
        k = int(np.argmin(ax))
        if ax[k] < 1e-6:
            keep[k] = False
            lin = k
    with np.errstate(divide="ignore"):
        if sinh_type:
            log_f = ax + np.log1p(-np.exp(-2.0 * ax))      # ln |2 sinh x|
        else:
            log_f = ax + np.log1p(np.exp(-2.0 * ax))       # ln 2 cosh x
    xs = x[keep]
    t = 1.0 / np.tanh(xs) if sinh_type else np.tanh(xs)  # F'/F
    a = h1[keep] * t
    b = h1[keep] ** 2 + h2[keep] * t                       # F''/F = 1 for cosh and sinh
    q1 = float(a.sum())
    q2 = float(b.sum()) + float(np.sum(a * (q1 - a)))      # + 2 sum_{r<s} a_r a_s
    log_scale = float(log_f[keep].sum())
    sign = float(np.prod(np.sign(xs))) if sinh_type else 1.0
    if lin is None:
        return log_scale, sign, 1.0, q1, q2
    # P = f Q with f = 2 sinh(x_k): P' = f' Q + f Q', P'' = f'' Q + 2 f' Q' + f Q''
    xk, hk1, hk2 = x[lin], h1[lin], h2[lin]
    f0 = 2.0 * math.sinh(xk)
    f1 = 2.0 * math.cosh(xk) * hk1
    f2 = f0 * hk1 * hk1 + 2.0 * math.cosh(xk) * hk2
    return log_scale, sign, f0, f1 + f0 * q1, f2 + 2.0 * f1 * q1 + f0 * q2


def _torus(K: float, m: int, n: int) -> tuple[float, float, float]:
    """(ln Z, d ln Z / dK, d2 ln Z / dK2) of the m x n torus."""
    N = m * n
    s = math.sinh(2.0 * K)
    pre = -math.log(2.0) + 0.5 * N * math.log(2.0 * s)
    pre1 = N * math.cosh(2.0 * K) / s
    pre2 = -2.0 * N / (s * s)
    terms = []
    for odd in (True, False):
        g, g1, g2 = _gamma(K, n, odd)
        for sinh_type in (False, True):
            terms.append(_product(m, g, g1, g2, sinh_type))
    top = max(t[0] for t in terms if np.isfinite(t[0]))
    z = z1 = z2 = 0.0
    for log_scale, sign, v0, v1, v2 in terms:
        if not np.isfinite(log_scale):
            continue
        w = sign * math.exp(log_scale - top)
        z += w * v0
        z1 += w * v1
        z2 += w * v2
    m1 = z1 / z
    return pre + top + math.log(z), pre1 + m1, pre2 + z2 / z - m1 * m1


def log_partition_torus(K: float, m: int, n: int) -> float:
    """ln Z of the m x n periodic square lattice at coupling K = 1/T."""
    return _torus(float(K), int(m), int(n))[0]


def energy_torus(T: float, L: int) -> float:
    """Exact internal energy per spin of the L x L torus, -(d ln Z / dK) / L^2."""
    return -_torus(1.0 / float(T), int(L), int(L))[1] / (L * L)


def specific_heat_torus(T: float, L: int) -> float:
    """Exact specific heat per spin of the L x L torus, K^2 (d2 ln Z / dK2) / L^2."""
    K = 1.0 / float(T)
    return K * K * _torus(K, int(L), int(L))[2] / (L * L)
