import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")
    config.addinivalue_line("markers", "ref: needs the compiled reference (oracle/_ref)")


@pytest.fixture(scope="session")
def gpu():
    """GPU tests fail (not skip) when no device is present: they are only
    selected with -m gpu, on a GPU machine."""
    import pseudoctimaging_amd as pct
    n = pct.device_count()
    assert n >= 1, "no HIP device visible"
    return n


@pytest.fixture(scope="session")
def ref():
    import oracle
    if not oracle.ref_available():
        pytest.skip("compiled reference (oracle/_ref) not available")
    return oracle
