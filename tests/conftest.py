import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library through the C-ABI)")


@pytest.fixture(autouse=True)
def _name_case(request):
    # name every case before it runs, so a run that dies on the GPU says where
    print(f"\n[case] {request.node.nodeid}", flush=True)
    yield
