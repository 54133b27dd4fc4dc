"""The allocation-free observables entry point (ising_msc_obs_static), through the C ABI: two
calls in a row, the second on a smaller lattice, each equal to ising_msc_obs with a caller
scratch (and to the oracle's exact observables of the same lattice)."""

import ctypes as C

import numpy as np
import pytest

from oracle import ising_oracle as O


@pytest.mark.gpu
def test_static_scratch_observables_equal_the_scratch_entry_point():
    from pyising_amd import _lib
    from pyising_amd.backends.hip_backend import OBS_SCRATCH_BYTES, HipMultiSpin

    L = _lib.lib()
    scratch = _lib.DeviceBuffer(OBS_SCRATCH_BYTES)
    out_a, out_b = _lib.DeviceBuffer(16), _lib.DeviceBuffer(16)
    # the larger lattice first (more reduction workgroups), then a smaller one: the second call
    # must not pick up partial sums the first left in the library's array
    for shape, seed in (((2048, 4096), 3), ((256, 384), 4)):
        print(f"no-alloc observables {shape}", flush=True)
        sim = HipMultiSpin(shape, 0.44, seed=seed)
        sim.sweep(3)
        for out in (out_a, out_b):
            _lib.check(L.ising_fill_u64(out.ptr, 2, 0x5A5A5A5A5A5A5A5A, None), "fill")
        _lib.check(L.ising_msc_obs(sim.w[0].ptr, sim.w[1].ptr, sim.rows, sim.nw, scratch.ptr,
                                   out_a.ptr, None), "msc_obs")
        _lib.check(L.ising_msc_obs_static(sim.w[0].ptr, sim.w[1].ptr, sim.rows, sim.nw,
                                          out_b.ptr, None), "msc_obs_static")
        a, b = np.zeros(2, np.int64), np.zeros(2, np.int64)
        out_a.download(a)
        out_b.download(b)
        assert a.tolist() == b.tolist()
        assert tuple(b.tolist()) == O.observables_exact(sim.get_lattice())
        sim.close()
    rc = L.ising_msc_obs_static(None, None, 16, 1, out_b.ptr, None)
    assert rc == -1 and b"null" in L.ising_last_error()
    rc = L.ising_msc_obs_static(out_a.ptr, out_a.ptr, 16, 1, None, None)
    assert rc == -1 and b"null" in L.ising_last_error()
    for h in (scratch, out_a, out_b):
        h.free()
