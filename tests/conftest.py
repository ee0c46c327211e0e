import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X) and the built libjgpm.so")


@pytest.fixture(scope="session")
def ctx():
    import jupiter_amd as J
    c = J.Context(0)
    print(f"[gpu] device {c.info()}", flush=True)
    yield c
    c.close()
