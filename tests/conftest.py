import importlib.util
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
TESTS = os.path.dirname(os.path.abspath(__file__))
if TESTS not in sys.path:
    sys.path.insert(0, TESTS)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")
    config.addinivalue_line("markers", "slow: long-running CPU test")


def load_hyg():
    """Import the package directory hyteg-vara_amd (its name is not a Python identifier)."""
    name = "hyteg_vara_amd"
    if name in sys.modules:
        return sys.modules[name]
    spec = importlib.util.spec_from_file_location(name, os.path.join(ROOT, "hyteg-vara_amd", "__init__.py"))
    mod = importlib.util.module_from_spec(spec)
    sys.modules[name] = mod
    spec.loader.exec_module(mod)
    return mod


@pytest.fixture(scope="session")
def hyg():
    return load_hyg()


@pytest.fixture(scope="session")
def hip_stream():
    """A HIP stream (non-default) for calls that need one, e.g. graph capture."""
    import ctypes as C
    lib = C.CDLL("libamdhip64.so")
    s = C.c_void_p()
    assert lib.hipStreamCreate(C.byref(s)) == 0
    yield s
    lib.hipStreamDestroy(s)
