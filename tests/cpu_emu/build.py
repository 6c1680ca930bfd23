"""Builds the CPU emulation of the engine library (tests only):
tests/cpu_emu/_build/librpd_engine_emu.so from rpd-statmc-denoiser_amd/csrc
compiled with g++ against include/hip/hip_runtime.h here, with
AddressSanitizer and UBSan bounds checks.

Each `extern __shared__ T name[];` (dynamic LDS) is rewritten into a pointer
to the emulated block's dynamic LDS; nothing else in the sources changes.
Run `python tests/cpu_emu/build.py`; load the result with libasan preloaded
(see tests/cpu_emu/run_emulated.py).
"""
import os
import re
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
SRC = os.path.join(REPO, "rpd-statmc-denoiser_amd", "csrc")
OUT = os.path.join(HERE, "_build")
LIB = os.path.join(OUT, "librpd_engine_emu.so")
DYN_LDS = re.compile(r"extern\s+__shared__\s+(\w+)\s+(\w+)\s*\[\s*\]\s*;")
FLAGS = ["-std=c++17", "-O1", "-g", "-fPIC", "-ffp-contract=off", "-fsanitize=address",
         "-fsanitize=bounds", "-fno-sanitize-recover=bounds", "-fno-omit-frame-pointer",
         "-Wno-unused-result", "-pthread"]


def prepare():
    src_dir = os.path.join(OUT, "src", "csrc")
    inc_dir = os.path.join(OUT, "src", "include")   # csrc/../../include from the copied tree
    shutil.rmtree(os.path.join(OUT, "src"), ignore_errors=True)
    os.makedirs(src_dir)
    os.makedirs(os.path.join(OUT, "include"), exist_ok=True)
    shutil.copy(os.path.join(REPO, "include", "rpd_engine.h"), os.path.join(OUT, "include", "rpd_engine.h"))
    units = []
    for name in sorted(os.listdir(SRC)):
        if name in REPLACED:
            continue
        text = open(os.path.join(SRC, name)).read()
        text = DYN_LDS.sub(r"\1* \2 = static_cast<\1*>(::rpd_emu::dynamic_lds());", text)
        if name.endswith(".hip"):
            name = name[:-4] + ".cpp"
            units.append(os.path.join(src_dir, name))
        with open(os.path.join(src_dir, name), "w") as f:
            f.write(text)
    return units


# engine sources that use device libraries the emulation has no g++ build of,
# each replaced by a CPU file of the same interface
REPLACED = {"sort_keys.hip": "sort_keys_emu.cpp"}


def compile_unit(cpp):
    obj = os.path.join(OUT, os.path.basename(cpp)[:-4] + ".o")
    cmd = ["g++", *FLAGS, "-I", os.path.join(HERE, "include"), "-c", cpp, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode:
        sys.stderr.write(r.stderr)
        raise SystemExit(f"failed: {' '.join(cmd)}")
    return obj


def build():
    units = prepare()
    units.append(os.path.join(HERE, "emu_runtime.cpp"))
    units += [os.path.join(HERE, f) for f in REPLACED.values()]
    with ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 1)) as ex:
        objs = list(ex.map(compile_unit, units))
    subprocess.run(["g++", "-shared", "-fsanitize=address", "-fsanitize=bounds", "-pthread", "-o", LIB, *objs, "-ldl"], check=True)
    return LIB


if __name__ == "__main__":
    print(build())
