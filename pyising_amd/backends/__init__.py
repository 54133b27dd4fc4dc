"""Backend registry of the reference's plugin API (src/pyising/backends/__init__.py:11-65).

``REGISTRY[(backend, variant)] = "module:Class"``, modules resolved under this package and
imported lazily; :func:`load` returns ``None`` when an entry cannot run here (no GPU, library
not built), :func:`create` raises ``RuntimeError`` then, and an unknown entry raises
``KeyError``.
"""

from __future__ import annotations

import importlib
from functools import cache

from .base import IsingSimulation

REGISTRY: dict[tuple[str, str], str] = {
    ("hip", "int8"): "hip_backend:HipInt8",
    ("hip", "msc"): "hip_backend:HipMultiSpin",
    ("hip", "msc-eager32"): "hip_backend:HipMultiSpinEager",
}

BACKENDS = ("hip",)
GPU_ONLY = {"hip"}


@cache
def load(backend: str, variant: str = "int8") -> type[IsingSimulation] | None:
    """The simulation class, or ``None`` if it cannot be used on this machine."""
    key = (backend, variant)
    if key not in REGISTRY:
        raise KeyError(f"unknown backend/variant {backend}/{variant}; known: {sorted(REGISTRY)}")
    if backend in GPU_ONLY:
        from .._lib import device_count

        if device_count() < 1:
            return None
    mod_name, cls_name = REGISTRY[key].split(":")
    try:
        mod = importlib.import_module(f"{__name__}.{mod_name}")
    except Exception:  # noqa: BLE001 - any import failure just disables the entry
        return None
    return getattr(mod, cls_name, None)


def available() -> list[tuple[str, str]]:
    """All registry entries that can run on this machine."""
    return [k for k in REGISTRY if load(*k) is not None]


def create(backend: str, variant: str = "int8", **kwargs) -> IsingSimulation:
    """Instantiate a simulation; raises ``RuntimeError`` if the entry is unavailable."""
    cls = load(backend, variant)
    if cls is None:
        raise RuntimeError(f"backend {backend}/{variant} is not available on this machine")
    return cls(**kwargs)


__all__ = ["REGISTRY", "BACKENDS", "IsingSimulation", "load", "available", "create"]
