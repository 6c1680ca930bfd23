"""The engine's kernels compiled for the CPU (tests/cpu_emu) under
AddressSanitizer and bounds checks, on a small render job and the golden
fixtures: a regression test for device-memory mistakes in the common path,
such as the unallocated sampling tape of render jobs that a dangling else in
job_alloc once caused (k_tape wrote through a null pointer).

Runs the GPU-marked tests named below in a subprocess with libasan preloaded
(tests/cpu_emu/run_emulated.py); the emulated library is rebuilt when any
engine source is newer than it.
"""
import glob
import os
import shutil
import subprocess
import sys

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
EMU = os.path.join(HERE, "cpu_emu")
LIB = os.path.join(EMU, "_build", "librpd_engine_emu.so")


def _stale():
    if not os.path.exists(LIB):
        return True
    srcs = glob.glob(os.path.join(REPO, "rpd-statmc-denoiser_amd", "csrc", "*")) + \
        glob.glob(os.path.join(EMU, "*.cpp")) + glob.glob(os.path.join(EMU, "include", "hip", "*.h")) + \
        [os.path.join(REPO, "include", "rpd_engine.h")]
    return max(os.path.getmtime(p) for p in srcs) > os.path.getmtime(LIB)


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++ with AddressSanitizer")
def test_render_job_and_golden_fixtures_under_asan():
    cmd = [sys.executable, os.path.join(EMU, "run_emulated.py")]
    if _stale():
        cmd.append("--rebuild")
    cmd += [os.path.join(HERE, "test_gpu_parity.py"), os.path.join(HERE, "test_golden.py"),
            "-k", "(test_trace_records_match_oracle and 0.0) or test_engine_against_golden_fixtures",
            "-q", "-p", "no:faulthandler"]
    r = subprocess.run(cmd, cwd=REPO, capture_output=True, text=True, timeout=1200)
    tail = "\n".join((r.stdout + r.stderr).splitlines()[-40:])
    assert r.returncode == 0, tail
    assert "2 passed" in r.stdout, tail
