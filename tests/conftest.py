import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X) and the built HIP runtime")


@pytest.fixture(autouse=True)
def _clear_fft_cache():
    yield
    try:
        import vkdispatch_amd as vd
        vd.fft.cache_clear()
    except Exception:
        pass
