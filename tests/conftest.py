import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and the built liblts.so")


@pytest.fixture(scope="session")
def gpu_ctx():
    from lighttransportsimulation_amd import Context
    ctx = Context(0)
    yield ctx
    ctx.close()
