import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
TESTS = Path(__file__).resolve().parent
for _p in (ROOT, TESTS):
    if str(_p) not in sys.path:
        sys.path.insert(0, str(_p))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X)")


@pytest.fixture(scope="session")
def cuda():
    """A HIP device seen through the library's own runtime (no torch)."""
    import pgdg_p2_amd as P
    if P.device_count() < 1:
        pytest.skip("no HIP device")
    return 0


@pytest.fixture(scope="session", autouse=True)
def oracle_threads():
    """The oracle's per-cell and per-particle loops on worker threads (its
    results do not depend on the count)."""
    import ctypes as C
    import os
    try:
        from oracle.pyoracle import lib
        L = lib()
    except Exception:
        yield
        return
    L.oracle_set_threads.argtypes = [C.c_int]
    n = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (os.cpu_count() or 1)
    L.oracle_set_threads(max(1, min(n, 16)))
    yield
    L.oracle_set_threads(1)
