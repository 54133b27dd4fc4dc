"""Exact results and statistical physics checks of the oracle (CPU only).

Mirrors tests/test_exact.py and tests/test_physics.py of the reference: Kaufman's torus
partition function against brute-force enumeration of the 4 x 4 torus, Onsager values, the
finite torus approaching the infinite lattice, and the oracle's equilibrium energy against the
exact torus energy within statistical errors.
"""

import math

import numpy as np
import pytest

from oracle import c_oracle as CO
from pyising_amd.exact import (
    TC_2D,
    energy_2d,
    energy_torus,
    log_partition_torus,
    magnetization_2d,
    specific_heat_torus,
)


def brute_force_4x4(T):
    confs = ((np.arange(2**16)[:, None] >> np.arange(16)) & 1).astype(np.int64) * 2 - 1
    confs = confs.reshape(-1, 4, 4)
    E = -(confs * np.roll(confs, 1, 1)).sum((1, 2)) - (confs * np.roll(confs, 1, 2)).sum((1, 2))
    w = np.exp(-(E - E.min()) / T)
    e = (E * w).sum() / w.sum() / 16
    e2 = (E * E * w).sum() / w.sum() / 256
    return np.log(w.sum()) - E.min() / T, e, (e2 - e * e) * 16 / T**2


@pytest.mark.parametrize("T", [1.0, 2.0, TC_2D, 3.0, 5.0])
def test_kaufman_matches_enumeration(T):
    lnz, e, c = brute_force_4x4(T)
    assert log_partition_torus(1 / T, 4, 4) == pytest.approx(lnz, rel=1e-12)
    assert energy_torus(T, 4) == pytest.approx(e, abs=1e-8)
    assert specific_heat_torus(T, 4) == pytest.approx(c, rel=1e-4)


def test_onsager_values():
    assert TC_2D == pytest.approx(2.269185314213022, rel=1e-14)
    assert energy_2d(TC_2D) == pytest.approx(-math.sqrt(2), rel=1e-9)
    assert magnetization_2d(3.0) == 0.0
    assert magnetization_2d(2.0) == pytest.approx((1 - math.sinh(1.0) ** -4) ** 0.125)


@pytest.mark.parametrize("T", [1.8, 2.6, 3.5])
def test_torus_approaches_infinite_lattice(T):
    assert energy_torus(T, 256) == pytest.approx(energy_2d(T), abs=1e-6)


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
