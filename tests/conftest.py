import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
TESTS = os.path.dirname(os.path.abspath(__file__))
if TESTS not in sys.path:
    sys.path.insert(0, TESTS)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and the built libgrscore.so")


@pytest.fixture(scope="session")
def engine():
    """One engine context per test session (GPU tests only)."""
    from grainrecon_amd import Engine

    e = Engine(0)
    yield e
    e.close()
