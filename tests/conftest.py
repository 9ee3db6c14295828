import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, 'oracle'))


def pytest_configure(config):
    config.addinivalue_line('markers', 'gpu: needs an AMD GPU (MI355X); runs the HIP engine through its C ABI')


@pytest.fixture(scope='session')
def engine():
    """One engine for the GPU tests (libpt through ctypes). Fails loudly when the library is not built."""
    from shadytracer_amd import pt
    e = pt.Engine()
    yield e
    e.close()
