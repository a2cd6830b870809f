import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "raytracer-gpupro4_amd")
for p in (PKG, os.path.join(ROOT, "oracle"), ROOT):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X) and the HIP library")
    config.addinivalue_line("markers", "slow: long-running CPU test")


@pytest.fixture(scope="session")
def rtcore():
    import rtcore as m

    m.build_library()
    return m


@pytest.fixture(scope="session")
def oracle():
    import oracle as m

    m.build_library()
    return m


@pytest.fixture(scope="session")
def ctx(rtcore):
    c = rtcore.Context(0)
    yield c
    c.close()
