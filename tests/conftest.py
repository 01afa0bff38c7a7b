import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (os.path.join(ROOT, "finetools.jl_amd"), os.path.join(ROOT, "oracle"), ROOT):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the C ABI on the device)")


@pytest.fixture(scope="session")
def ctx():
    import feasm

    c = feasm.Context(0)
    yield c
    c.close()


@pytest.fixture(scope="session")
def hiprt():
    """The HIP runtime libfeasm links (the system ROCm), loaded with ctypes
    for device buffers of the tests of the device-pointer entry points."""
    import ctypes
    import ctypes.util

    for name in ("libamdhip64.so", "/opt/rocm/lib/libamdhip64.so", ctypes.util.find_library("amdhip64")):
        if not name:
            continue
        try:
            h = ctypes.CDLL(name)
        except OSError:
            continue
        h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        h.hipFree.argtypes = [ctypes.c_void_p]
        return h
    pytest.skip("HIP runtime library not found")
