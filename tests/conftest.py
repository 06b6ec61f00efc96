import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for sub in ("axiom-physics-engine_amd", "oracle", "tests"):
    p = os.path.join(ROOT, sub)
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path through the C ABI)")


def gpu_available():
    """True when libaxm sees a HIP device (the library's own runtime; no torch in-process)."""
    try:
        import axm as m
        return m.device_count() > 0
    except Exception:
        return False


@pytest.fixture(scope="session")
def axm():
    import axm as m
    m.lib()
    return m


@pytest.fixture(scope="session")
def orc():
    import orc as o
    o.lib()
    return o


@pytest.fixture(scope="session")
def gpu():
    if not gpu_available():
        pytest.skip("no GPU")
    return True
