import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for sub in ("splatkit-ios_amd", "oracle", "tests"):
    path = os.path.join(ROOT, sub)
    if path not in sys.path:
        sys.path.insert(0, path)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


@pytest.fixture(scope="session")
def oracle():
    import pyoracle

    pyoracle.build()
    return pyoracle


@pytest.fixture(scope="session")
def ref(oracle):
    r = oracle.ref()
    if r is None:
        pytest.skip("oracle/_ref/libref.so not built (no reference tree on this machine)")
    return r


@pytest.fixture(scope="session")
def skm_mod():
    import skm

    return skm


@pytest.fixture(scope="session")
def renderer(skm_mod):
    """One renderer for the GPU session; every test sets its own extent, policy and world."""
    r = skm_mod.Renderer()
    print(f"device: {r.description()}", flush=True)
    yield r
    r.close()
