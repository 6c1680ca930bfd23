import os
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(REPO, "tests"))
sys.path.insert(0, os.path.join(REPO, "rpd-statmc-denoiser_amd"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP engine)")


def _gpu_available():
    try:
        import rpd_engine
        n = __import__("ctypes").c_int()
        # hipGetDeviceCount through the engine: creating a context fails without a device
        ctx = rpd_engine.Context(0)
        ctx.close()
        return True
    except Exception:
        return False


@pytest.fixture(scope="session")
def engine_ctx():
    import rpd_engine
    ctx = rpd_engine.Context(0)
    yield ctx
    ctx.close()
