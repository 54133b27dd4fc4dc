"""Host side of the drop-in boundary, without a GPU: the registry and error behaviour of the
reference's plugin API (src/pyising/backends/__init__.py:33-62, base.py:44-73), shape
validation, seeds, hot-start draws, and the serving/batching host logic."""

import numpy as np
import pytest

import pyising_amd as P
from oracle import ising_oracle as O
from pyising_amd import backends as B
from pyising_amd import lattice as Lt


def test_registry_entries_and_order():
    assert ("hip", "int8") in B.REGISTRY and ("hip", "msc") in B.REGISTRY
    assert ("hip", "msc-eager32") in B.REGISTRY
    assert B.BACKENDS == ("hip",)


def test_unknown_entry_is_key_error():
    with pytest.raises(KeyError):
        B.load("hip", "nope")
    with pytest.raises(KeyError):
        P.create("cuda", "int8", shape=(16, 16), beta=0.4)


def test_unavailable_entry_is_runtime_error_without_gpu():
    from pyising_amd import _lib

    if _lib.device_count() > 0:
        pytest.skip("a GPU is visible")
    assert B.load("hip", "msc") is None
    assert B.available() == []
    with pytest.raises(RuntimeError):
        P.create("hip", "msc", shape=(128, 128), beta=0.4)


@pytest.mark.parametrize("variant,shape", [("int8", (15, 16)), ("int8", (16, 12)),
                                           ("int8", (8,)), ("msc", (16, 64)),
                                           ("msc", (8, 8, 128))])
def test_shape_validation_precedes_device_use(variant, shape):
    from pyising_amd.backends.hip_backend import HipInt8, HipMultiSpin

    cls = HipInt8 if variant == "int8" else HipMultiSpin
    with pytest.raises(ValueError):
        cls(shape, 0.4)


def test_algorithm_and_init_validation():
    from pyising_amd.backends.hip_backend import HipInt8, HipMultiSpin

    with pytest.raises(ValueError):
        HipMultiSpin((128, 128), 0.4, algorithm="heatbath")
    with pytest.raises(ValueError):
        HipInt8((16, 16), 0.4, algorithm="glauber")


def test_check_shape_matches_reference_rules():
    assert Lt.check_shape((16, 24)) == (16, 24)
    assert Lt.check_shape([4, 6, 8]) == (4, 6, 8)
    for bad in [(8,), (7, 8), (8, 12), (0, 8), (8, 8, 8, 8)]:
        with pytest.raises(ValueError):
            Lt.check_shape(bad, 8)
    with pytest.raises(ValueError):
        Lt.check_shape((16, 192), 128)


def test_split_seed_matches_oracle():
    for seed in (0, 1, 7, 2**40 + 3, 2**64 - 1, -5):
        assert Lt.split_seed(seed) == O.split_seed(seed)


def test_constants():
    assert Lt.TC_2D == pytest.approx(2.269185314213022, rel=1e-14)
    assert np.sinh(2 / Lt.TC_2D) == pytest.approx(1.0, rel=1e-14)
    assert Lt.BETA_C_3D == 0.221654626


def test_row_parity_matches_oracle():
    for shape in [(8, 16), (4, 6, 8)]:
        np.testing.assert_array_equal(Lt.row_parity(shape), O.row_parity(shape))
