import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a real MI355X (runs through the C ABI)")


@pytest.fixture(scope="session")
def ctx():
    import torch
    from crava_amd.core import Context
    assert torch.cuda.is_available(), "GPU test needs a GPU"
    c = Context()
    yield c
    c.close()
