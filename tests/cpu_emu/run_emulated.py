"""Runs GPU-marked tests against the CPU emulation of the engine's kernels
(tests only): builds tests/cpu_emu/_build/librpd_engine_emu.so if needed and
starts pytest with libasan preloaded and RPD_CPU_EMU=1.

    python tests/cpu_emu/run_emulated.py tests/test_metrics.py -k device -x
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, HERE)
import build  # noqa: E402


def main(argv):
    if not os.path.exists(build.LIB) or "--rebuild" in argv:
        build.build()
    argv = [a for a in argv if a != "--rebuild"]
    asan = subprocess.run(["g++", "-print-file-name=libasan.so"], capture_output=True, text=True).stdout.strip()
    env = dict(os.environ, RPD_CPU_EMU="1", RPD_CPU_EMU_LIB=build.LIB, LD_PRELOAD=asan,
               ASAN_OPTIONS="detect_leaks=0:abort_on_error=0:halt_on_error=1:allocator_may_return_null=1")
    cmd = [sys.executable, "-m", "pytest", "-m", "gpu", "-p", "no:cacheprovider", *argv]
    return subprocess.call(cmd, cwd=REPO, env=env)


if __name__ == "__main__":
    sys.exit(main(sys.argv[1:]))
