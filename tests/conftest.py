import os
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (REPO, os.path.join(REPO, "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and the built libobvhs_hip.so")


@pytest.fixture(scope="session")
def ctx():
    import obvhs_amd
    c = obvhs_amd.Context(0)
    c.set_trace(True)
    yield c
    c.close()
