import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path)")


@pytest.fixture(scope="session")
def oracle():
    from oracle.bind import Oracle
    return Oracle()


@pytest.fixture(scope="session")
def oracle64():
    from oracle.bind import Oracle
    return Oracle(fp64=True)


@pytest.fixture(scope="session")
def ref_headers():
    from oracle.bind import RefHeaders
    if not os.path.exists(RefHeaders.PATH):
        pytest.skip("oracle/_ref not built: the reference sources are not mounted here")
    return RefHeaders()


@pytest.fixture(scope="session")
def qlib():
    from qtrk_amd import api
    return api.lib()


@pytest.fixture(scope="session")
def gpu(qlib):
    """The HIP library on a real GPU; fails (never skips) when the device is missing."""
    import ctypes as C
    buf = C.create_string_buffer(512)
    if qlib.QTrkGetDeviceInfo(buf, 512) != 0:
        raise RuntimeError("no HIP device: " + qlib.QTrkGetLastError().decode())
    print("device:", buf.value.decode(), flush=True)
    return qlib
