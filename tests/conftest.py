import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and the built library")
    config.addinivalue_line("markers", "slow: long CPU oracle runs")


@pytest.fixture(scope="session")
def gpu_ctx():
    """One library context for the whole GPU session."""
    import ligesis_amd

    ctx = ligesis_amd.Context(0)  # raises when no HIP device or library is present
    yield ctx
    ctx.close()
