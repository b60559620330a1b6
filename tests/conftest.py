"""Shared test set-up: import paths, markers, and the oracle build.

`gpu` tests need a visible MI355X and the built libshyhip.so; everything
else runs on the CPU.  The CPU oracle (oracle/) is the checker only."""
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "shyfemcm-ismar_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU and the built libshyhip.so")


@pytest.fixture(scope="session", autouse=True)
def _oracle_built():
    import pyoracle
    for prec in ("f32", "f64"):
        if not os.path.exists(os.path.join(ROOT, "oracle", f"liboracle_{prec}.so")):
            pyoracle.build()
            break
    yield


@pytest.fixture(scope="session")
def gpu():
    import shyhip
    if not shyhip.gpu_available():
        pytest.skip("no GPU / libshyhip.so not built")
    return shyhip
