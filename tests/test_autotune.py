"""Batch autotuner: host bookkeeping on the CPU; on the GPU, every tuned choice (per-shape msc
tile, launch form, graph length) leaves every lattice bit-identical to sweeping it alone."""

import numpy as np
import pytest

from oracle import ising_oracle as O
from pyising_amd.autotune import BatchTuning, signature


class _S:
    def __init__(self, variant, shape, algorithm="metropolis"):
        self.variant, self.shape, self.algorithm = variant, shape, algorithm


def test_signature_is_order_independent_and_counts():
    a = [_S("int8", (64, 64)), _S("msc", (128, 256)), _S("int8", (64, 64), "heatbath"),
         _S("int8", (64, 64))]
    b = list(reversed(a))
    assert signature(a) == signature(b)
    assert "int8:64x64:metropolis*2" in signature(a)
    assert signature(a) != signature(a[:3])


def test_tuning_json_round_trip(tmp_path, monkeypatch):
    from pyising_amd import autotune

    t = BatchTuning(msc_tiles={(1024, 1024): "lockstep", (128, 256): "queue"},
                    int8_vecs={(64, 64): 8, (8, 8, 16): 4}, persistent=True,
                    graph_sweeps=10, tiles_per_workgroup=4, flips_per_ns=123.4)
    assert BatchTuning.from_json(t.to_json()) == t
    monkeypatch.setenv("PYISING_AMD_TUNE_CACHE", str(tmp_path / "tune.json"))
    autotune._save_cache({"k": t.to_json()})
    assert BatchTuning.from_json(autotune._load_cache()["k"]) == t


def _mixed(hip):
    specs = [(hip.HipInt8, (64, 64), 1 / 2.2, "metropolis"),
             (hip.HipInt8, (8, 8, 16), 0.22, "heatbath"),
             (hip.HipMultiSpin, (128, 256), 0.44, "metropolis"),
             (hip.HipMultiSpin, (256, 8192), 1 / 2.269, "metropolis"),
             (hip.HipMultiSpin, (64, 384), 1 / 2.25, "metropolis")]
    return [(cls, shape, beta, alg, 300 + i, O.random_lattice(shape, 40 + i))
            for i, (cls, shape, beta, alg) in enumerate(specs)]


@pytest.fixture(scope="module")
def hip():
    from pyising_amd.backends import hip_backend

    return hip_backend


@pytest.mark.gpu
@pytest.mark.parametrize("tiles", [{(128, 256): "lockstep", (256, 8192): "queue"},
                                   {(128, 256): "queue", (256, 8192): "lockstep",
                                    (64, 384): "lockstep"}])
def test_per_shape_msc_tiles_are_exact(hip, tiles):
    from pyising_amd.batch import BatchScan

    specs = _mixed(hip)
    sims = [cls(sh, b, seed=sd, algorithm=al, init=lat) for cls, sh, b, al, sd, lat in specs]
    refs = [cls(sh, b, seed=sd, algorithm=al, init=lat) for cls, sh, b, al, sd, lat in specs]
    scan = BatchScan(sims, msc_tiles=tiles)
    scan.sweep(6)
    scan.synchronize()
    for sim, ref in zip(sims, refs, strict=True):
        ref.sweep(6)
        np.testing.assert_array_equal(sim.get_lattice(), ref.get_lattice())
    scan.close()


def test_vec_choices():
    from pyising_amd.autotune import _vec_choices

    assert _vec_choices((64, 64)) == [16, 8, 4]
    assert _vec_choices((8, 8, 16)) == [8, 4]
    assert _vec_choices((16, 24)) == [4]


@pytest.mark.gpu
@pytest.mark.parametrize("vecs", [{(64, 64): 4, (8, 8, 16): 8}, {(64, 64): 8, (8, 8, 16): 4},
                                  {(64, 64): 16}])
def test_per_shape_int8_widths_are_exact(hip, vecs):
    from pyising_amd.batch import BatchScan

    specs = _mixed(hip)
    sims = [cls(sh, b, seed=sd, algorithm=al, init=lat) for cls, sh, b, al, sd, lat in specs]
    refs = [cls(sh, b, seed=sd, algorithm=al, init=lat) for cls, sh, b, al, sd, lat in specs]
    scan = BatchScan(sims, int8_vecs=vecs)
    scan.sweep(6)
    scan.synchronize()
    for sim, ref in zip(sims, refs, strict=True):
        ref.sweep(6)
        np.testing.assert_array_equal(sim.get_lattice(), ref.get_lattice())
    scan.close()


@pytest.mark.gpu
def test_bad_int8_width_is_rejected(hip):
    from pyising_amd import _lib
    from pyising_amd.batch import BatchScan

    sims = [hip.HipInt8((16, 24), 0.4, seed=1)]  # Lx/2 = 12: only 4 sites per lane divide it
    with pytest.raises(_lib.HipError, match="vec"):
        BatchScan(sims, int8_vecs={(16, 24): 8})


@pytest.mark.gpu
def test_tuned_scan_is_exact(hip, tmp_path, monkeypatch):
    from pyising_amd.batch import BatchScan

    monkeypatch.setenv("PYISING_AMD_TUNE_CACHE", str(tmp_path / "tune.json"))
    specs = _mixed(hip)
    sims = [cls(sh, b, seed=sd, algorithm=al, init=lat) for cls, sh, b, al, sd, lat in specs]
    refs = [cls(sh, b, seed=sd, algorithm=al, init=lat) for cls, sh, b, al, sd, lat in specs]
    scan = BatchScan.tuned(sims, sweeps=10)
    t = scan.tuning
    assert t.flips_per_ns > 0 and len(t.trials) >= 3
    assert set(t.msc_tiles) == {(128, 256), (256, 8192), (64, 384)}
    n = 0
    if scan.graph_sweeps:
        scan.replay(2)
        n += 2 * scan.graph_sweeps
    scan.sweep(3)
    n += 3
    scan.synchronize()
    for sim, ref in zip(sims, refs, strict=True):
        ref.sweep(n)
        assert sim.t == ref.t == n
        np.testing.assert_array_equal(sim.get_lattice(), ref.get_lattice())
    scan.close()
    # second call: the cache answers without re-timing
    again = BatchScan.tuned(sims, sweeps=10)
    assert again.tuning.to_json() == t.to_json()
    again.close()
