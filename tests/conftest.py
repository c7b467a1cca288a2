import os
import sys

import pytest

# torch ships its own HIP runtime under the same soname as ROCm's; whichever
# loads first serves the whole process.  Load torch's before libfasthip so
# the tests that hand torch buffers to the library share one runtime, as the
# benchmark does.
try:
    import torch  # noqa: F401
except ImportError:
    torch = None

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


@pytest.fixture(scope="session")
def oracle():
    from oracle import pyoracle
    pyoracle.lib()
    return pyoracle


@pytest.fixture(scope="session")
def fh():
    import fastc_amd
    fastc_amd.lib()
    if not fastc_amd.device_available():
        pytest.fail("GPU test selected but no HIP device is available")
    return fastc_amd
