"""pytest setup: the `gpu` marker, import paths, and shared fixtures."""
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "gaussian-splat-renderer_amd")
for p in (ROOT, PKG, os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels through the C ABI)")


@pytest.fixture(scope="session")
def oracle():
    import oracle as o  # noqa: F401  (oracle/oracle.py, the CPU checker)

    return o


@pytest.fixture(scope="session")
def gsr_ctx():
    import gsr

    ctx = gsr.Context(0)
    yield ctx
    ctx.close()
