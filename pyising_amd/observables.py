"""Measurement statistics on the host: time series, binning errors, the Binder cumulant, its
jackknife error and the crossing of two Binder curves (src/pyising/observables.py:1-85; same
names, arguments and formulas).

Beside the reference's time-series interface, :func:`binder_from_blocks` and
:func:`scan_binder` work from per-block moment sums: a batched scan accumulates sum M^2 and
sum M^4 on the device (exact integers), so the block means, and from them the Binder cumulant
and its blocked jackknife error, need no per-sweep host transfer.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np


@dataclass
class Series:
    """Magnetisation ``m`` and energy ``e`` per spin, one entry per measurement."""

    m: np.ndarray
    e: np.ndarray

    def summary(self) -> dict:
        """Means with binning errors, <m^2>, <m^4>, Binder cumulant with jackknife error
        (src/pyising/observables.py:17-26)."""
        m = np.asarray(self.m, dtype=float)
        e = np.asarray(self.e, dtype=float)
        abs_m = np.abs(m)
        return {"abs_m": float(abs_m.mean()), "abs_m_err": binning_error(abs_m),
                "e": float(e.mean()), "e_err": binning_error(e),
                "m2": float(np.mean(m * m)), "m4": float(np.mean(m**4)),
                "binder": binder(m), "binder_err": jackknife_binder_error(m),
                "n": int(m.size)}


def _blocks(x: np.ndarray, n_bins: int) -> np.ndarray | None:
    """The leading ``nb * (size // nb)`` samples as ``nb`` equal blocks (rows), or None when
    fewer than two blocks are possible."""
    x = np.asarray(x, dtype=float).ravel()
    nb = min(int(n_bins), x.size)
    if nb < 2:
        return None
    per = x.size // nb
    return x[: nb * per].reshape(nb, per)


def binning_error(x: np.ndarray, n_bins: int = 32) -> float:
    """Standard error of the mean from ``n_bins`` block means (src/pyising/observables.py:29-37);
    robust to autocorrelation once blocks are longer than the autocorrelation time."""
    b = _blocks(x, n_bins)
    if b is None:
        return float("nan")
    means = b.mean(axis=1)
    return float(means.std(ddof=1) / np.sqrt(means.size))


def binder(m: np.ndarray) -> float:
    """U_4 = 1 - <m^4> / (3 <m^2>^2) (src/pyising/observables.py:40-43)."""
    m = np.asarray(m, dtype=float)
    m2 = np.mean(m * m)
    return float(1.0 - np.mean(m**4) / (3.0 * m2 * m2))


def binder_from_blocks(m2: np.ndarray, m4: np.ndarray) -> tuple[float, float]:
    """Binder cumulant and blocked jackknife error from per-block means of m^2 and m^4 (equal
    block lengths). Leave-one-block-out estimates U_i use the other blocks' means; the error is
    sqrt((nb - 1) * mean((U_i - mean U)^2)) (the estimator of
    src/pyising/observables.py:46-59)."""
    m2 = np.asarray(m2, dtype=float)
    m4 = np.asarray(m4, dtype=float)
    nb = m2.size
    u = 1.0 - m4.mean() / (3.0 * m2.mean() ** 2)
    if nb < 2:
        return float(u), float("nan")
    loo2 = (m2.sum() - m2) / (nb - 1)
    loo4 = (m4.sum() - m4) / (nb - 1)
    jk = 1.0 - loo4 / (3.0 * loo2 * loo2)
    return float(u), float(np.sqrt((nb - 1) * np.mean((jk - jk.mean()) ** 2)))


def jackknife_binder_error(m: np.ndarray, n_bins: int = 32) -> float:
    """Blocked jackknife error of :func:`binder` (src/pyising/observables.py:46-59)."""
    b = _blocks(m, n_bins)
    if b is None:
        return float("nan")
    return binder_from_blocks((b * b).mean(axis=1), (b**4).mean(axis=1))[1]


def sample(sim, n_therm: int, n_meas: int, every: int = 1) -> Series:
    """Thermalise for ``n_therm`` sweeps, then take ``n_meas`` measurements ``every`` sweeps
    apart (src/pyising/observables.py:62-75). Any object with ``sweep(n)`` and either
    ``observables()`` or ``magnetization()``/``energy()`` works; one with
    ``measure_series(n, every)`` (the hip simulations: exact (M, E) pairs gathered on the
    device) is asked for the whole series at once."""
    sim.sweep(int(n_therm))
    series = getattr(sim, "measure_series", None)
    if series is not None:  # device-side series: same numbers, no host wait per measurement
        me = series(int(n_meas), int(every)).astype(float) / sim.n_spins
        return Series(me[:, 0].copy(), me[:, 1].copy())
    m = np.empty(int(n_meas))
    e = np.empty(int(n_meas))
    measure = getattr(sim, "observables", None)
    for i in range(int(n_meas)):
        sim.sweep(int(every))
        if measure is not None:
            m[i], e[i] = measure()
        else:
            m[i], e[i] = sim.magnetization(), sim.energy()
    return Series(m, e)


def crossing(Ts: np.ndarray, u_a: np.ndarray, u_b: np.ndarray) -> float:
    """First temperature where two Binder curves cross: the first zero of u_a - u_b on the
    grid, else linear interpolation inside the first sign change; NaN if none
    (src/pyising/observables.py:78-85)."""
    Ts = np.asarray(Ts, dtype=float)
    d = np.asarray(u_a, dtype=float) - np.asarray(u_b, dtype=float)
    for i in range(d.size - 1):
        if d[i] == 0.0:
            return float(Ts[i])
        if d[i] * d[i + 1] < 0.0:
            return float(Ts[i] + (Ts[i + 1] - Ts[i]) * d[i] / (d[i] - d[i + 1]))
    return float("nan")


def scan_binder(scan, n_blocks: int, sweeps_per_block: int) -> list[tuple[float, float]]:
    """Binder cumulant and jackknife error of every lattice of a batched scan
    (:class:`pyising_amd.batch.BatchScan`) from ``n_blocks`` blocks of ``sweeps_per_block``
    sweeps: the device sums of each block give that block's means of m^2 and m^4."""
    n_lat = len(scan.sims)
    m2 = np.zeros((n_lat, n_blocks))
    m4 = np.zeros((n_lat, n_blocks))
    for b in range(int(n_blocks)):
        scan.reset_moments()
        scan.sweep(int(sweeps_per_block))
        for i, mo in enumerate(scan.moments()):
            m2[i, b] = mo.mean(2)
            m4[i, b] = mo.mean(4)
    return [binder_from_blocks(m2[i], m4[i]) for i in range(n_lat)]
