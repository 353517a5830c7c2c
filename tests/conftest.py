import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


@pytest.fixture(scope="session")
def gpu():
    """GPU tests fail (not skip) when no device is visible: they are only
    collected with -m gpu, which is run on a GPU machine."""
    import fsmeld_amd
    n = fsmeld_amd.device_count()
    assert n >= 1, "no HIP device visible to a -m gpu test"
    return n
