"""GPU test cases held off the GPU.

Each file marked with UNEXPLAINED_FAULT was running (alone or with others)
when a GPU memory fault or hung queue ended the whole run, with no output
left to name the failing test, and reading the code has not shown a cause.
The device code these tests launch has since been made instruction-identical
to a build on which they passed repeatedly; that removes differences without
naming the one that mattered, so it does not count as a fix, and the tests
stay skipped until a cause is shown.
"""
import os

import pytest

# Under the CPU emulation of the kernels (tests/cpu_emu, RPD_CPU_EMU=1) these
# tests run: nothing of them reaches a GPU there.
UNEXPLAINED_FAULT = pytest.mark.skipif(
    os.environ.get("RPD_CPU_EMU") != "1",
    reason="held off the GPU: a run of these tests ended in a GPU fault whose cause is not identified")
