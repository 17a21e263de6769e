import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and libfeddgpu.so")


@pytest.fixture(scope="session")
def gpu_ctx():
    from feddlib_amd import native
    if native.device_count() < 1:
        pytest.fail("no HIP device visible for a gpu-marked test")
    ctx = native.Context(0)
    yield ctx
    ctx.close()
