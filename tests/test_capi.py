"""The C-ABI library loads without a GPU and exports every symbol include/ising_hip.h declares,
with the argument counts the ctypes layer uses (no compute calls here)."""

import re
from pathlib import Path

import pytest

from pyising_amd import _lib

ROOT = Path(__file__).resolve().parents[1]
HEADER = ROOT / "include" / "ising_hip.h"


def declared() -> dict[str, int]:
    text = re.sub(r"/\*.*?\*/", "", HEADER.read_text(), flags=re.S)
    out = {}
    for m in re.finditer(r"\b(?:int|const char\*)\s+(ising_\w+)\s*\(([^;]*?)\)\s*;", text, re.S):
        args = m.group(2).strip()
        out[m.group(1)] = 0 if args in ("", "void") else args.count(",") + 1
    return out


def test_header_parses():
    decl = declared()
    assert len(decl) >= 30
    assert "ising_msc_half" in decl and "ising_int8_half" in decl


def test_library_exports_every_declared_symbol():
    lib = _lib.load_library()
    for name in declared():
        assert hasattr(lib, name), name


def test_ctypes_signatures_match_header():
    decl = declared()
    for name, nargs in decl.items():
        if name == "ising_last_error":
            assert nargs == 0
            continue
        assert name in _lib.SIGNATURES, f"{name} declared but not bound"
        assert len(_lib.SIGNATURES[name]) == nargs, name
    assert set(_lib.SIGNATURES) <= set(decl)


def test_errors_without_gpu_are_loud():
    if _lib.device_count() > 0:
        pytest.skip("a GPU is visible")
    with pytest.raises(_lib.HipError):
        _lib.lib()
    from pyising_amd.backends.hip_backend import HipMultiSpin

    with pytest.raises(_lib.HipError):
        HipMultiSpin((128, 128), 0.4)


def test_argument_validation_needs_no_gpu():
    """Bad arguments are rejected by the C layer before any HIP call."""
    lib = _lib.load_library()
    rc = lib.ising_int8_half(None, None, 16, 12, 16, 1, 2, 0, None, 0, 0, 0, 0, None)
    assert rc == -1
    assert b"null" in lib.ising_last_error()
    rc = lib.ising_msc_config(7, 8)
    assert rc == -1 and b"mode" in lib.ising_last_error()
    for k in (43, 72, 22, 41, 151):
        rc = lib.ising_msc_half_ex(1, 1, 16, 1, 0, 0, 0, 0, 0, 0, 0, k, None, None)
        assert rc == -1 and b"cap" in lib.ising_last_error()
    rc = lib.ising_msc_half_ex(1, 1, 16, 1, 0, 0, 0, 0, 0, 0, 5, 42, None, None)
    assert rc == -1 and b"mode" in lib.ising_last_error()
    for k in (44, 62, 22):  # the fused form has 512-word tiles (R = 2) and cap 3..5
        rc = lib.ising_msc_half_ex(1, 1, 16, 1, 0, 0, 0, 0, 0, 0, 4, k, None, None)
        assert rc == -1 and b"fused" in lib.ising_last_error()
    rc = lib.ising_msc_sweeps_sync(1, 1, 16, 1, 0, 1, 0, 0, 0, 0, 4, 42, None, None, None, None)
    assert rc == -1 and b"workspace" in lib.ising_last_error()
    rc = lib.ising_msc_half_ex(1, 1, 15, 1, 0, 0, 0, 0, 0, 0, 0, 42, None, None)
    assert rc == -1 and b"geometry" in lib.ising_last_error()
    rc = lib.ising_split_pack(1, 1, 1, 4, 100, None)
    assert rc == -1 and b"128" in lib.ising_last_error()


def test_msc_kernel_by_size_needs_no_gpu():
    """The default msc kernel is the lockstep tile up to 2^18 words per colour (e.g. 4096 x
    8192) and queue compaction above; explicit modes pass through."""
    import ctypes as C

    lib = _lib.load_library()
    mode, k = C.c_int(), C.c_int()

    def kernel(rows, nw, m=-1, kk=42):
        assert lib.ising_msc_kernel(rows, nw, m, kk, C.byref(mode), C.byref(k)) == 0
        return mode.value, k.value

    assert kernel(1024, 8) == (1, 0)
    assert kernel(4096, 64) == (1, 0)  # 2^18 words
    assert kernel(4098, 64) == (0, 42)
    assert kernel(8192, 64) == (0, 42)  # 2^19 words
    assert kernel(16384, 128) == (0, 42)
    assert kernel(16384, 128, 3, 44) == (0, 44)
    assert kernel(1024, 8, 0, 32) == (0, 32)
    assert kernel(16384, 128, 1, 0) == (1, 0)
    assert kernel(16384, 128, 2, 0) == (2, 0)
    assert lib.ising_msc_kernel(1024, 8, 5, 42, C.byref(mode), C.byref(k)) == -1
    assert lib.ising_msc_kernel(1024, 8, 3, 41, C.byref(mode), C.byref(k)) == -1


def test_msc_process_default_needs_no_gpu():
    """ising_msc_config sets the process default that mode -1 resolves to; modes 1 and 2 keep
    the compaction k set before, which mode 0 / 3 then use again."""
    import ctypes as C

    lib = _lib.load_library()
    mode, k = C.c_int(), C.c_int()

    def default(rows, nw):
        assert lib.ising_msc_kernel(rows, nw, -1, 0, C.byref(mode), C.byref(k)) == 0
        return mode.value, k.value

    try:
        assert lib.ising_msc_config(0, 52) == 0
        assert default(1024, 8) == (0, 52)
        assert lib.ising_msc_config(1, 0) == 0
        assert default(16384, 128) == (1, 0)
        assert lib.ising_msc_config(0, 99) == -1  # rejected: the default is unchanged
        assert default(16384, 128) == (1, 0)
        assert lib.ising_msc_config(3, 32) == 0
        assert default(1024, 8) == (1, 0) and default(16384, 128) == (0, 32)
    finally:
        assert lib.ising_msc_config(3, 42) == 0
    assert default(16384, 128) == (0, 42)


def test_int8_kernel_choice_needs_no_gpu():
    """Sites per lane and wave layout of the int8 per-half-sweep kernel, host only: segment
    waves when a row's chunks divide 64 and those rows divide the plane, row waves when a row
    holds a multiple of 64 chunks, per-lane rows otherwise."""
    import ctypes as C

    lib = _lib.load_library()

    def choice(shape):
        ndim = len(shape)
        lz = shape[0] if ndim == 3 else 1
        v, lay = C.c_int(), C.c_int()
        assert lib.ising_int8_kernel(shape[-2] * lz, shape[-1] // 2, shape[-2], lz, ndim,
                                     C.byref(v), C.byref(lay)) == 0
        return v.value, lay.value

    assert choice((8192, 8192)) == (16, 1)      # 256 chunks per row
    assert choice((256, 256, 256)) == (16, 2)   # 8 chunks per row, 8 rows per wave
    assert choice((64, 64)) == (4, 2)
    assert choice((256, 4096)) == (4, 1)
    assert choice((16, 16, 16)) == (4, 0)       # 32 rows per wave > 16 rows per plane
    assert choice((10, 12, 16)) == (4, 0)       # 240 chunks: not whole waves
    v = C.c_int()
    assert lib.ising_int8_kernel(10, 5, 10, 1, 2, C.byref(v), C.byref(v)) == -1  # lh % 4

