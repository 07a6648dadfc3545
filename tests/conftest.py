import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the C ABI on HIP)")


@pytest.fixture(scope="session")
def ctx():
    import zim_amd
    return zim_amd.Context(0)


@pytest.fixture(scope="session", params=["tiled", "plain"])
def kctx(request):
    import zim_amd
    c = zim_amd.Context(0)
    c.set_kernels(request.param)
    return c
