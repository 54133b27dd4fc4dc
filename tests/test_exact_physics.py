"""Exact results (pyising_amd/exact.py) checked by independent means, and the oracle's physics
against them (CPU only).

* The torus partition function and its K-derivatives against brute-force enumeration of small
  rectangular tori (m != n, both orientations) and against a row-to-row transfer matrix of a
  6 x 6 torus (Z = Tr T^6, energy from Tr(T^5 T') with T' = dT/dK, specific heat from the
  transfer-matrix ln Z by a second difference).
* The elliptic integral (arithmetic-geometric mean) against Gauss-Legendre quadrature of its
  defining integral, and Onsager's energy against the torus energy of a large torus (two
  independent formulas for the same limit).
* The oracle's equilibrium energy against the exact torus energy, within statistical errors.
"""

import itertools
import math

import numpy as np
import pytest

from oracle import c_oracle as CO
from pyising_amd.exact import (
    TC_2D,
    elliptic_k,
    energy_2d,
    energy_torus,
    log_partition_torus,
    magnetization_2d,
    specific_heat_torus,
)
from pyising_amd.exact import _torus


def enumerate_torus(T, m, n):
    """ln Z, energy per spin and specific heat per spin of the m x n torus by summing over
    all 2^(mn) configurations."""
    N = m * n
    conf = ((np.arange(2**N)[:, None] >> np.arange(N)) & 1).astype(np.int8) * 2 - 1
    c = conf.reshape(-1, m, n).astype(np.int64)
    E = -(c * np.roll(c, 1, 1)).sum((1, 2)) - (c * np.roll(c, 1, 2)).sum((1, 2))
    e0 = E.min()
    w = np.exp(-(E - e0) / T)
    Z = w.sum()
    mean = (E * w).sum() / Z
    var = ((E - mean) ** 2 * w).sum() / Z
    return math.log(Z) - e0 / T, mean / N, var / N / T**2


def transfer_matrix_torus(K, L):
    """ln Z and -d ln Z/dK of the L x L torus from the symmetric row-to-row transfer matrix
    T[s, t] = exp(K (h(s)/2 + h(t)/2 + v(s, t))), h = bonds within a row, v = bonds between
    rows: Z = Tr T^L, dZ/dK = L Tr(T^(L-1) T') with T' = T * (h/2 + h/2 + v)."""
    rows = np.array(list(itertools.product((-1, 1), repeat=L)), dtype=np.int64)
    h = (rows * np.roll(rows, 1, axis=1)).sum(1)
    v = rows @ rows.T
    B = 0.5 * h[:, None] + 0.5 * h[None, :] + v
    # the largest weight scaled out of every factor, so that the powers stay finite
    T = np.exp(K * B - K * B.max())
    TL1 = np.linalg.matrix_power(T, L - 1)
    Z = np.trace(TL1 @ T)
    dZ = L * np.trace(TL1 @ (T * B))
    return math.log(Z) + L * K * B.max(), dZ / Z


@pytest.mark.parametrize("m,n", [(3, 5), (5, 3), (4, 5), (2, 7)])
@pytest.mark.parametrize("T", [1.2, 2.0, TC_2D, 3.0, 6.0])
def test_torus_against_enumeration(m, n, T):
    lnz, e, c = enumerate_torus(T, m, n)
    l, d1, d2 = _torus(1.0 / T, m, n)
    assert l == pytest.approx(lnz, rel=1e-12)
    assert -d1 / (m * n) == pytest.approx(e, abs=1e-10)
    assert d2 / (m * n) / T**2 == pytest.approx(c, rel=1e-9)


@pytest.mark.parametrize("T", [1.5, TC_2D, 2.8, 4.0])
def test_torus_against_transfer_matrix(T):
    L = 6
    K = 1.0 / T
    lnz, d1 = transfer_matrix_torus(K, L)
    assert log_partition_torus(K, L, L) == pytest.approx(lnz, rel=1e-12)
    assert energy_torus(T, L) == pytest.approx(-d1 / (L * L), abs=1e-10)
    h = 1e-4
    second = (transfer_matrix_torus(K + h, L)[0] - 2 * lnz + transfer_matrix_torus(K - h, L)[0])
    assert specific_heat_torus(T, L) == pytest.approx(K * K * second / h**2 / (L * L), rel=1e-5)


@pytest.mark.parametrize("kappa", [0.0, 0.3, 0.9, 0.999, 0.999999])
def test_elliptic_k_against_quadrature(kappa):
    # substitution theta = pi/2 (1 - u^2) removes the near-singularity at theta = pi/2
    x, w = np.polynomial.legendre.leggauss(400)
    u = 0.5 * (x + 1.0)
    theta = 0.5 * np.pi * (1.0 - u * u)
    integrand = np.pi * u / np.sqrt(1.0 - (kappa * np.sin(theta)) ** 2)
    assert float(elliptic_k(kappa)) == pytest.approx(0.5 * float(np.sum(w * integrand)), rel=1e-9)


def test_onsager_and_yang_values():
    assert TC_2D == pytest.approx(2.269185314213022, rel=1e-14)
    assert math.sinh(2.0 / TC_2D) == pytest.approx(1.0, rel=1e-14)
    assert energy_2d(TC_2D) == pytest.approx(-math.sqrt(2.0), rel=1e-12)
    assert magnetization_2d(3.0) == 0.0 and magnetization_2d(TC_2D) == 0.0
    assert magnetization_2d(2.0) == pytest.approx((1 - math.sinh(1.0) ** -4) ** 0.125, rel=1e-14)
    # ground state and infinite temperature limits
    assert energy_2d(0.05) == pytest.approx(-2.0, abs=1e-12)
    assert energy_2d(1e6) == pytest.approx(0.0, abs=1e-5)


@pytest.mark.parametrize("T", [1.6, 2.1, 2.5, 3.5])
def test_large_torus_approaches_onsager(T):
    """Two independent formulas for the same limit: the 512 x 512 torus energy (Kaufman,
    analytic derivative) and Onsager's infinite-lattice energy (elliptic integral)."""
    assert energy_torus(T, 512) == pytest.approx(energy_2d(T), abs=1e-9)


def test_specific_heat_peaks_near_tc():
    Ts = np.linspace(2.1, 2.5, 81)
    c = [specific_heat_torus(T, 32) for T in Ts]
    t_peak = Ts[int(np.argmax(c))]
    assert TC_2D - 0.02 < t_peak < TC_2D + 0.05  # finite-size peak just above T_c


def _blocked_mean_err(x, nb=20):
    b = np.asarray(x[: len(x) // nb * nb]).reshape(nb, -1).mean(axis=1)
    return b.mean(), b.std(ddof=1) / np.sqrt(nb)


@pytest.mark.parametrize("variant", ["int8-metropolis", "int8-heatbath", "msc"])
@pytest.mark.parametrize("T", [1.8, 3.0])
def test_oracle_energy_matches_exact_torus(variant, T):
    """The oracle's equilibrium energy per spin equals the exact torus energy within 5 sigma."""
    if variant == "msc":
        L = 128
        sim = CO.CMsc((L, L), 1 / T, seed=3, init="cold")
    else:
        L = 16
        sim = CO.CInt8((L, L), 1 / T, seed=3, algorithm=variant.split("-")[1], init="cold")
    sim.sweep(300)
    es = []
    for _ in range(2000 if L == 16 else 600):
        sim.sweep(2)
        es.append(sim.observables_exact()[1] / (L * L))
    e, err = _blocked_mean_err(es)
    assert abs(e - energy_torus(T, L)) < 5 * err + 2e-3
