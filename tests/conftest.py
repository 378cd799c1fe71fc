import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path)")


def gpu_available():
    """A HIP device is visible.  Asked of the HIP runtime directly: torch's
    answer turns False once another library in the process has initialised
    HIP first, which silently skipped later GPU tests."""
    try:
        from heimdall_multibeam_amd import hip
        return hip.device_count() > 0
    except Exception:
        return False


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle as O
    O.build()
    return O
