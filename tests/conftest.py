import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU and the built HIP library")
    config.addinivalue_line("markers", "slow: long-running CPU check")


def _has_gpu() -> bool:
    try:
        from pyising_amd import _lib

        return _lib.device_count() > 0
    except Exception:  # noqa: BLE001
        return False


def pytest_collection_modifyitems(config, items):
    if _has_gpu():
        return
    skip = pytest.mark.skip(reason="needs a HIP GPU")
    for item in items:
        if "gpu" in item.keywords:
            item.add_marker(skip)


@pytest.fixture(autouse=True, scope="session")
def _session_tune_cache(tmp_path_factory):
    """The batch autotuner caches its choices in a file; tests use a fresh one of their own,
    never the user's cache."""
    import os

    key = "PYISING_AMD_TUNE_CACHE"
    old = os.environ.get(key)
    os.environ[key] = str(tmp_path_factory.mktemp("tune") / "batch_tune.json")
    yield
    if old is None:
        os.environ.pop(key, None)
    else:
        os.environ[key] = old
