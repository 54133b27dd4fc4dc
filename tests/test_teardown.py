"""Release of HIP handles at interpreter shutdown (CPU only, with a recording stand-in for the
library): once shutdown has begun, garbage-collected handles call nothing in the library;
explicit close()/free() calls release at any time before it."""

import subprocess
import sys
import textwrap
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]


class Recorder:
    """Stand-in for the ctypes library: records every call by name."""

    def __init__(self):
        self.calls = []

    def __getattr__(self, name):
        if not name.startswith("ising_"):
            raise AttributeError(name)

        def fn(*args):
            self.calls.append(name)
            return 0

        return fn


def _handles(rec):
    from pyising_amd import _lib
    from pyising_amd.batch import BatchScan

    buf = object.__new__(_lib.DeviceBuffer)
    buf._lib, buf.ptr, buf.nbytes = rec, 0x1000, 64
    ev = object.__new__(_lib.Event)
    ev._lib, ev.ptr = rec, 0x2000
    st = object.__new__(_lib.Stream)
    st._lib, st.ptr = rec, 0x3000
    scan = object.__new__(BatchScan)
    scan._L, scan._h = rec, 0x4000
    scan.stream = type("S", (), {"sync": lambda self: rec.calls.append("stream_sync")})()
    return [buf, ev, st, scan]


def test_finalizers_release_before_shutdown(monkeypatch):
    from pyising_amd import _lib

    monkeypatch.setattr(_lib, "_finalizing", False)
    rec = Recorder()
    for h in _handles(rec):
        h.__del__()
    assert {"ising_free", "ising_event_destroy", "ising_stream_destroy",
            "ising_batch_destroy"} <= set(rec.calls)


def test_finalizers_call_nothing_once_shutdown_began(monkeypatch):
    from pyising_amd import _lib

    rec = Recorder()
    hs = _handles(rec)
    monkeypatch.setattr(_lib, "_finalizing", True)
    assert not _lib.releasing_allowed()
    for h in hs:
        h.__del__()
    assert rec.calls == []


def test_explicit_close_releases_once(monkeypatch):
    from pyising_amd import _lib

    monkeypatch.setattr(_lib, "_finalizing", False)
    rec = Recorder()
    buf, ev, st, scan = _handles(rec)
    buf.free()
    ev.close()
    st.close()
    scan.close()
    n = len(rec.calls)
    for h in (buf, ev, st, scan):  # a second release and the finalizer do nothing
        h.__del__()
    buf.free()
    ev.close()
    st.close()
    assert len(rec.calls) == n


def test_no_library_call_during_real_interpreter_shutdown():
    """A handle still referenced by a module global when the interpreter exits (collected
    during finalization) must not reach the library."""
    code = textwrap.dedent(f"""
        import sys
        sys.path.insert(0, {str(ROOT)!r})
        from pyising_amd import _lib

        class Rec:
            def __getattr__(self, name):
                def fn(*a):
                    sys.stderr.write("LIBCALL " + name + "\\n")
                    sys.stderr.flush()
                    return 0
                return fn

        keep = object.__new__(_lib.DeviceBuffer)
        keep._lib, keep.ptr, keep.nbytes = Rec(), 0x1000, 8
        ev = object.__new__(_lib.Event)
        ev._lib, ev.ptr = Rec(), 0x2000
        cycle = [ev]
        cycle.append(cycle)  # collected by the final garbage collection
        print("done", flush=True)
    """)
    res = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True,
                         timeout=120)
    assert res.returncode == 0, res.stderr
    assert "done" in res.stdout
    assert "LIBCALL" not in res.stderr


@pytest.mark.parametrize("name", ["bench.py"])
def test_bench_releases_explicitly(name):
    """The benchmark closes what it created before returning (no reliance on finalizers)."""
    text = (ROOT / name).read_text()
    assert ".close()" in text


def test_bench_calls_only_existing_handle_methods():
    """Every method bench.py calls on its HIP handles (events, streams, device buffers) exists
    on the handle classes: the GPU path of bench.py runs on no CPU test, so a misspelt call
    would first show on the GPU."""
    import ast

    from pyising_amd import _lib

    tree = ast.parse((ROOT / "bench.py").read_text())
    handles = {}  # variable name -> class, from `x = _lib.Event()` style assignments
    for node in ast.walk(tree):
        if isinstance(node, ast.Assign) and isinstance(node.value, (ast.Call, ast.Tuple)):
            values = node.value.elts if isinstance(node.value, ast.Tuple) else [node.value]
            targets = node.targets[0].elts if isinstance(node.targets[0], ast.Tuple) else [node.targets[0]]
            for t, v in zip(targets, values):
                if (isinstance(t, ast.Name) and isinstance(v, ast.Call)
                        and isinstance(v.func, ast.Attribute) and isinstance(v.func.value, ast.Name)
                        and v.func.value.id == "_lib" and v.func.attr in ("Event", "Stream", "DeviceBuffer")):
                    handles[t.id] = getattr(_lib, v.func.attr)
    assert {"e0", "e1"} <= set(handles), handles
    called = [(n.func.value.id, n.func.attr) for n in ast.walk(tree)
              if isinstance(n, ast.Call) and isinstance(n.func, ast.Attribute)
              and isinstance(n.func.value, ast.Name) and n.func.value.id in handles]
    assert called
    missing = [(v, m) for v, m in called if not hasattr(handles[v], m)]
    assert not missing, missing
