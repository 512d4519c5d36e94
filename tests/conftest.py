import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
sys.path.insert(0, str(ROOT / "cms-julia_amd"))
sys.path.insert(0, str(ROOT / "tests"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


@pytest.fixture(scope="session")
def frames():
    from pixeltrack import generator as G
    return G.frames()


def gpu_available():
    try:
        import ctypes as C
        from pixeltrack import _abi as A
        n = C.c_int(0)
        return A.hip_lib().pt_device_count(C.byref(n)) == 0 and n.value > 0
    except Exception:
        return False
