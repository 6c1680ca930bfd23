import os
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(REPO, "tests"))
sys.path.insert(0, os.path.join(REPO, "rpd-statmc-denoiser_amd"))

# RPD_CPU_EMU=1 (set by tests/cpu_emu/run_emulated.py only): the GPU tests run
# against the engine's kernels compiled for the CPU under AddressSanitizer
# (tests/cpu_emu) instead of the HIP library. A debugging aid for kernel
# indexing and synchronisation; it is not the product and proves nothing
# about the GPU build.
if os.environ.get("RPD_CPU_EMU") == "1":
    import rpd_engine as _eng
    _eng.LIB_PATH = os.environ.get("RPD_CPU_EMU_LIB",
                                   os.path.join(REPO, "tests", "cpu_emu", "_build", "librpd_engine_emu.so"))


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
