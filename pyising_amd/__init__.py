"""pyising_amd: an MI355X-native ``hip`` backend for pyising's checkerboard Ising Monte Carlo.

Same plugin API as the reference (src/pyising/__init__.py:3-9):

>>> from pyising_amd import create, TC_2D
>>> sim = create("hip", "msc", shape=(16384, 16384), beta=1 / TC_2D, seed=1)
>>> sim.sweep(100)
>>> m, e = sim.observables()
"""

from .backends import BACKENDS, REGISTRY, IsingSimulation, available, create, load
from .lattice import BETA_C_2D, BETA_C_3D, TC_2D, TC_3D

__all__ = ["create", "load", "available", "REGISTRY", "BACKENDS", "IsingSimulation", "TC_2D",
           "BETA_C_2D", "BETA_C_3D", "TC_3D"]
__version__ = "0.1.0"
