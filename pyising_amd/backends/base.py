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
    """One lattice and its sweep counter on one backend.

    The constructor takes the lattice extents (2D or 3D, each even, the last one a multiple of
    :attr:`layout_multiple`), the inverse temperature, the 64-bit key of the random stream, the
    update rule (Metropolis or heat bath) and the start: random (``"hot"``), all +1
    (``"cold"``) or a given int8 lattice of +1 / -1.
    """

    backend: ClassVar[str] = "base"
    variant: ClassVar[str] = "base"
    dims: ClassVar[tuple[int, ...]] = (2, 3)
    algorithms: ClassVar[tuple[str, ...]] = ("metropolis", "heatbath")
    layout_multiple: ClassVar[int] = 8
    #: which random stream the variant draws from; two simulations drawing from the same
    #: stream with the same key and start stay bit-identical
    stream: ClassVar[str] = "int8"
    #: the traffic model: bytes moved per attempted update (own word read and written, the
    #: other colour read once)
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
        """New inverse temperature from the next sweep on; spins and sweep counter stay."""
        self.beta = float(beta)
        self.table = threshold_table(self.beta, self.ndim, self.algorithm)
        self._upload_table()

    def sweep(self, n: int = 1) -> None:
        """``n`` sweeps; each updates colour 0, then colour 1, then advances the counter."""
        for _ in range(int(n)):
            self._half_sweep(0)
            self._half_sweep(1)
            self.t += 1

    @abstractmethod
    def _setup(self) -> None:
        """Allocate what the backend keeps besides the spins, which the start fills."""

    @abstractmethod
    def _half_sweep(self, color: int) -> None:
        """All sites of one colour, drawing from the stream at sweep counter ``self.t``."""

    @abstractmethod
    def get_lattice(self) -> np.ndarray:
        """A host copy of the lattice, int8 +1 / -1, in the constructor's shape."""

    @abstractmethod
    def set_lattice(self, lat: np.ndarray) -> None:
        """Replace the spins by a host lattice of the constructor's shape."""

    def _upload_table(self) -> None:  # noqa: B027 - optional hook
        """Hand a new threshold table to the device (the hip kernels take it as arguments,
        so nothing is uploaded)."""

    @abstractmethod
    def _fill_hot(self) -> None:
        """Hot start, bit-identical to numpy ``default_rng(seed)`` draws."""

    def _fill_cold(self) -> None:
        self.set_lattice(np.ones(self.shape, dtype=np.int8))

    def synchronize(self) -> None:  # noqa: B027 - optional hook
        """Return once the device has finished the simulation's queued work."""

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
        """Device bytes the spins occupy."""
        return int(self.n_spins * self.bytes_per_update / 3.0)
