import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through librbgpu's C ABI)")


@pytest.fixture(scope="session")
def ctx():
    from robast_amd._lib import Context
    c = Context(0)
    yield c
    c.close()
