import os
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (os.path.join(REPO, "tests"), os.path.join(REPO, "oracle"),
          os.path.join(REPO, "t2-relion_amd", "python")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libt2x on the GPU)")


@pytest.fixture(scope="session")
def orc():
    import oracle_py
    return oracle_py.load()


@pytest.fixture(scope="session")
def gpu_device():
    import ctypes as C
    import t2x_py as T
    lib = T.load()
    h = C.c_void_p()
    T.check(lib.t2x_device_open(0, C.byref(h)), "t2x_device_open")
    yield h
    lib.t2x_device_close(h)
