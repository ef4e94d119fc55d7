import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (os.path.join(ROOT, "accelerating-tc_amd"), os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libtcamd.so on cuda:0)")


@pytest.fixture(scope="session")
def ctx():
    import tcamd
    c = tcamd.Context(0)
    yield c
    c.close()
