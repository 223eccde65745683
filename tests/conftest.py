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
