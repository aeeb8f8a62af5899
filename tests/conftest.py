import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU; runs the HIP path")
    config.addinivalue_line("markers", "slow: full-size workload")


@pytest.fixture(scope="session")
def ctx():
    """One device context shared by the GPU tests (fails loudly without the library)."""
    import asmooth_amd as A
    c = A.Context(0)
    yield c
    c.close()
