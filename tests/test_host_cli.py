"""Host-side drop-in checks without a GPU: the config parser against the
reference's own config test (path_tracer/tests/config_test.cpp:5-56) and
every config file the reference ships (copied under tests/golden/configs),
and the command line's error paths (main.cpp:1597-1623, 1643-1836).
"""
import os
import subprocess
import tempfile

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(REPO, "rpd-statmc-denoiser_amd", "bin")
CONFIGS = os.path.join(REPO, "tests", "golden", "configs")

MINIMAL = """image_width: 16
image_height: 8
samples_per_pixel: 2
fov: 45
camera_position: [0, 0, 0]
camera_direction: [0, 0, 1]
scene: cornell_box
output: test
"""


def check(text):
    with tempfile.NamedTemporaryFile("w", suffix=".yaml", delete=False) as f:
        f.write(text)
        path = f.name
    try:
        r = subprocess.run([os.path.join(BIN, "config_check"), path], capture_output=True, text=True, timeout=30)
    finally:
        os.unlink(path)
    vals = dict(line.split(" ", 1) for line in r.stdout.strip().splitlines() if " " in line)
    return r.returncode, vals


def test_minimal_config_defaults():
    rc, v = check(MINIMAL)
    assert rc == 0, v
    assert v["camera_up"] == "0 1 0"
    assert v["tonemap"] == "agx"
    assert v["rpf_tile_size"] == "8"
    assert v["adaptive_importance_smoothing_radius"] == "1"
    assert v["focus_distance"] == "1"   # default = |camera_direction| (config.h:90)


def test_unknown_key_rejected():
    rc, v = check(MINIMAL + "unused_setting: 1\n")
    assert rc == 1 and "unused_setting" in v.get("error", "")


def test_invalid_value_rejected():
    rc, v = check(MINIMAL + "color_sigma_max: 0\n")
    assert rc == 1 and "color_sigma_max" in v.get("error", "")


@pytest.mark.parametrize("name", sorted(os.listdir(CONFIGS)))
def test_reference_config_files_parse(name):
    rc, v = check(open(os.path.join(CONFIGS, name)).read())
    assert rc == 0, v
    assert int(v["image_width"]) > 0 and v["scene"]


def _run(*args):
    return subprocess.run([os.path.join(BIN, "path_tracer"), *args], capture_output=True, text=True, timeout=60)


def test_cli_requires_config():
    r = _run()
    assert r.returncode == 106 and "--config is required" in (r.stdout + r.stderr)


def test_cli_rejects_both_statmc_flags():
    r = _run("-c", os.path.join(CONFIGS, "cornell.yaml"), "--statmc", "--no-statmc")
    assert r.returncode == 1 and "Cannot specify both --statmc and --no-statmc" in (r.stdout + r.stderr)


def test_cli_missing_config_file():
    r = _run("-c", "/nonexistent/config.yaml")
    assert r.returncode == 1
