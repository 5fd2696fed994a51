import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


def _have_gpu():
    # the kernel driver node exists only where a GPU is visible; on such a
    # host a missing or broken liberp.so must fail the tests, not skip them
    return os.path.exists("/dev/kfd")


@pytest.fixture(scope="session")
def gpu_ctx():
    """One HIP context for the whole GPU session (the library must be built)."""
    if not _have_gpu():
        pytest.skip("no GPU")
    from europaraypaths_amd._native import Context
    ctx = Context(0)
    yield ctx
    ctx.close()
