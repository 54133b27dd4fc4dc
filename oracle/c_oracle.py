"""ctypes wrapper of ``oracle/libising_oracle.so``, the C restatement of the oracle.

TEST INFRASTRUCTURE ONLY (tests, smoke check, benchmark CPU baseline). Pinned bit for bit to
:mod:`oracle.ising_oracle` by tests/test_oracle_c.py.
"""

from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

from . import ising_oracle as O

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "libising_oracle.so"

_P = C.c_void_p
_I64 = C.c_int64
_U32 = C.c_uint32
_INT = C.c_int

_SIGS = {
    "oracle_philox": [_P, _P, _P, _I64],
    "oracle_num_threads": [],
    "oracle_set_threads": [_INT],
    "oracle_int8_half": [_P, _P, _I64, _I64, _I64, _I64, _INT, _INT, _P, _U32, _U32, _U32, _U32],
    "oracle_int8_sweeps": [_P, _P, _I64, _I64, _I64, _I64, _INT, _INT, _P, _U32, _I64, _U32, _U32],
    "oracle_msc_half": [_P, _P, _I64, _I64, _U32, _U32, _U32, _U32, _U32, _U32, _P],
    "oracle_msc_sweeps": [_P, _P, _I64, _I64, _U32, _I64, _U32, _U32, _U32, _U32, _P],
    "oracle_int8_obs": [_P, _P, _I64, _I64, _I64, _I64, _INT, _P],
    "oracle_msc_obs": [_P, _P, _I64, _I64, _P],
}

_lib = None


def build() -> Path:
    """Compile the C restatement with the committed Makefile."""
    subprocess.run(["make", "-s", "-C", str(HERE), "-B", "libising_oracle.so"], check=True)
    return LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            build()
        _lib = C.CDLL(str(LIB_PATH))
        for name, args in _SIGS.items():
            f = getattr(_lib, name)
            f.argtypes = args
            f.restype = _INT if name == "oracle_num_threads" else None
    return _lib


def _ptr(a: np.ndarray):
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data


def philox(ctr: np.ndarray, key: np.ndarray) -> np.ndarray:
    ctr = np.ascontiguousarray(ctr, np.uint32).reshape(-1, 4)
    key = np.ascontiguousarray(key, np.uint32).reshape(-1, 2)
    out = np.empty_like(ctr)
    lib().oracle_philox(_ptr(ctr), _ptr(key), _ptr(out), ctr.shape[0])
    return out


def set_threads(n: int) -> None:
    lib().oracle_set_threads(int(n))


def num_threads() -> int:
    return int(lib().oracle_num_threads())


class CInt8:
    """int8 compressed-checkerboard simulation driven by the C restatement."""

    def __init__(self, shape, beta, seed=0, algorithm="metropolis", init="hot"):
        ref = O.Int8Oracle(shape, beta, seed=seed, algorithm=algorithm, init=init)
        self.shape, self.ndim, self.algorithm = ref.shape, ref.ndim, algorithm
        self.seed, self.t, self.rows, self.lh = ref.seed, 0, ref.rows, ref.lh
        self.ly = self.shape[-2]
        self.lz = self.shape[0] if self.ndim == 3 else 1
        self.table = ref.table
        self.c = [np.ascontiguousarray(x) for x in ref.c]

    def set_beta(self, beta):
        self.table = O.threshold_table(float(beta), self.ndim, self.algorithm)

    def sweep(self, n=1):
        k0, k1 = O.split_seed(self.seed)
        lib().oracle_int8_sweeps(_ptr(self.c[0]), _ptr(self.c[1]), self.rows, self.lh, self.ly,
                                 self.lz, self.ndim, int(self.algorithm == "heatbath"),
                                 _ptr(self.table), self.t & 0xFFFFFFFF, int(n), k0, k1)
        self.t += int(n)

    def get_lattice(self):
        return O.from_compressed(self.c[0], self.c[1], self.shape)

    def observables_exact(self):
        out = np.zeros(2, np.int64)
        lib().oracle_int8_obs(_ptr(self.c[0]), _ptr(self.c[1]), self.rows, self.lh, self.ly,
                              self.lz, self.ndim, _ptr(out))
        return int(out[0]), int(out[1])


class CMsc:
    """Multi-spin-coded 2D simulation driven by the C restatement."""

    def __init__(self, shape, beta, seed=0, algorithm="metropolis", init="hot", lattice=None):
        ref = O.MscOracle(shape, beta, seed=seed, init=init if lattice is None else lattice)
        self.shape, self.seed, self.t = ref.shape, ref.seed, 0
        self.rows, self.nw = ref.rows, ref.nw
        self.table = ref.table
        self.w = [np.ascontiguousarray(x) for x in ref.w]
        self.calls = np.zeros(1, np.uint64)

    def set_beta(self, beta):
        self.table = O.threshold_table(float(beta), 2, "metropolis")

    def sweep(self, n=1):
        k0, k1 = O.split_seed(self.seed)
        lib().oracle_msc_sweeps(_ptr(self.w[0]), _ptr(self.w[1]), self.rows, self.nw,
                                self.t & 0xFFFFFFFF, int(n), k0, k1, int(self.table[3]),
                                int(self.table[4]), _ptr(self.calls))
        self.t += int(n)

    def get_lattice(self):
        return O.from_compressed(O.unpack_bits(self.w[0]), O.unpack_bits(self.w[1]), self.shape)

    def observables_exact(self):
        out = np.zeros(2, np.int64)
        lib().oracle_msc_obs(_ptr(self.w[0]), _ptr(self.w[1]), self.rows, self.nw, _ptr(out))
        return int(out[0]), int(out[1])
