import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (os.path.join(ROOT, "rt-dsl_amd"), os.path.join(ROOT, "oracle"), ROOT):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the C ABI of libgprt_hip.so)")


@pytest.fixture(scope="session")
def ctx():
    import gprt_hip
    c = gprt_hip.Context(0)
    yield c
    c.close()
