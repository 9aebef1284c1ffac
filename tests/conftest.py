import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libdfrocm on the device)")


@pytest.fixture(scope="session")
def gpu_ctx():
    """The process's device context; fails loudly when the native library
    or the device is missing (no CPU fallback exists)."""
    from dataframe_mojo_amd import runtime

    return runtime.context(0)
