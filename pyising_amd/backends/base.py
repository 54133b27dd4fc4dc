"""The simulation contract of the reference's plugin API (src/pyising/backends/base.py:19-139).

Same constructor, hooks, ClassVars and error behaviour, so a ``hip`` simulation is a drop-in
for any reference backend: ``cls(shape, beta, seed=0, algorithm="metropolis", init="hot")``,
``sweep(n)``, ``set_beta``, ``get_lattice``/``set_lattice``, ``observables()``.
"""

from __future__ import annotations

from abc import ABC, abstractmethod
from typing import ClassVar

import numpy as np

from ..lattice import check_shape, threshold_table


class IsingSimulation(ABC):
    """A checkerboard Ising Monte Carlo simulation living on one backend.

    shape: ``(L, L)`` or ``(L, L, L)``; extents even, last extent a multiple of
    :attr:`layout_multiple`. beta: inverse temperature. seed: 64-bit Philox key.
    algorithm: ``"metropolis"`` or ``"heatbath"``. init: ``"hot"``, ``"cold"`` or an explicit
    int8 lattice.
    """

    backend: ClassVar[str] = "base"
    variant: ClassVar[str] = "base"
    dims: ClassVar[tuple[int, ...]] = (2, 3)
    algorithms: ClassVar[tuple[str, ...]] = ("metropolis", "heatbath")
    layout_multiple: ClassVar[int] = 8
    #: random-stream family: same stream, seed and start give bit-identical lattices
    stream: ClassVar[str] = "int8"
    #: minimal DRAM traffic per attempted spin update (read own + neighbours, write own)
    bytes_per_update: ClassVar[float] = 3.0
    gpu: ClassVar[bool] = True

    def __init__(self, shape, beta: float, seed: int = 0, algorithm: str = "metropolis",
                 init="hot") -> None:
        self.shape = check_shape(shape, self.layout_multiple)
        self.ndim = len(self.shape)
        if self.ndim not in self.dims:
            raise ValueError(f"{self.backend}/{self.variant} supports dims {self.dims}")
        if algorithm not in self.algorithms:
            raise ValueError(f"{self.backend}/{self.variant} supports {self.algorithms}")
        self.algorithm = algorithm
        self.seed = int(seed) & 0xFFFFFFFFFFFFFFFF
        self.beta = float(beta)
        self.t = 0
        self.rows = int(np.prod(self.shape[:-1]))
        self.lh = self.shape[-1] // 2
        self.table = threshold_table(self.beta, self.ndim, self.algorithm)
        self._setup()
        if isinstance(init, np.ndarray):
            self.set_lattice(init)
        elif init == "cold":
            self._fill_cold()
        elif init == "hot":
            self._fill_hot()
        else:
            raise ValueError(f"unknown init {init!r}")

    @property
    def n_spins(self) -> int:
        return int(np.prod(self.shape))

    def set_beta(self, beta: float) -> None:
        """Change the temperature; keeps the configuration and the sweep counter."""
        self.beta = float(beta)
        self.table = threshold_table(self.beta, self.ndim, self.algorithm)
        self._upload_table()

    def sweep(self, n: int = 1) -> None:
        """Run ``n`` full checkerboard sweeps (colour 0 then colour 1)."""
        for _ in range(int(n)):
            self._half_sweep(0)
            self._half_sweep(1)
            self.t += 1

    @abstractmethod
    def _setup(self) -> None:
        """Prepare device state (the spins are allocated by the init path)."""

    @abstractmethod
    def _half_sweep(self, color: int) -> None:
        """Update every site of one colour using sweep counter ``self.t``."""

    @abstractmethod
    def get_lattice(self) -> np.ndarray:
        """The configuration as a full int8 host array."""

    @abstractmethod
    def set_lattice(self, lat: np.ndarray) -> None:
        """Load a full int8 host configuration."""

    def _upload_table(self) -> None:  # noqa: B027 - optional hook
        """Push ``self.table`` to the device (tables travel as kernel arguments here)."""

    @abstractmethod
    def _fill_hot(self) -> None:
        """Hot start, bit-identical to numpy ``default_rng(seed)`` draws."""

    def _fill_cold(self) -> None:
        self.set_lattice(np.ones(self.shape, dtype=np.int8))

    def synchronize(self) -> None:  # noqa: B027 - optional hook
        """Block until all queued device work is done."""

    @abstractmethod
    def observables_exact(self) -> tuple[int, int]:
        """``(M, E)`` as exact integers (on-device integer reductions)."""

    def observables(self) -> tuple[float, float]:
        """``(m, e)`` per spin."""
        m, e = self.observables_exact()
        return m / self.n_spins, e / self.n_spins

    def magnetization(self) -> float:
        return self.observables()[0]

    def energy(self) -> float:
        return self.observables()[1]

    def device_bytes(self) -> int:
        """Bytes of lattice state held on the device."""
        return int(self.n_spins * self.bytes_per_update / 3.0)
