import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X) and the built HIP library")


@pytest.fixture(scope="session")
def gpu_lib():
    """The built library with a visible GPU; GPU tests fail (not skip) if the
    library is missing, so a broken build cannot pass silently."""
    from nutrlan_amd import trlan as T

    L = T.lib()
    assert T.device_count() > 0, "no GPU visible to the HIP runtime"
    return T
