import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "matter-compressor_amd")
for p in (PKG, os.path.join(ROOT, "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path)")


@pytest.fixture(scope="session")
def gpu_ctx():
    import mc_amd
    ctx = mc_amd.Ctx(0)
    yield ctx
    ctx.close()


@pytest.fixture(scope="session")
def kat():
    import numpy as np
    d = np.load(os.path.join(ROOT, "tests", "golden", "block_kat.npz"))
    return {k: d[k] for k in d.files}
