import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "ifrit-v2_amd"))
sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")


@pytest.fixture(scope="session")
def renderer_factory():
    from ifrit_hip import Renderer

    made = []

    def make():
        r = Renderer()
        made.append(r)
        return r

    yield make
    for r in made:
        r.close()
