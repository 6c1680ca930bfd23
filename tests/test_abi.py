"""The C ABI library loads without a GPU, exports every function the public
header declares, and its structs have the sizes the Python mirrors assume."""
import ctypes as C
import os
import re

import numpy as np

import rpd_engine as eng

HEADER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include", "rpd_engine.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", " ", text, flags=re.S)
    return sorted(set(re.findall(r"\b(rpd_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_function():
    lib = C.CDLL(eng.LIB_PATH)
    names = declared_functions()
    assert len(names) > 40
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_struct_sizes_match_python_mirrors():
    out = np.zeros(14, np.int32)
    eng.check(eng.lib().rpd_abi_struct_sizes(out.ctypes.data_as(C.c_void_p), 14))
    mirrors = [eng.RenderConfig, eng.Camera, eng.Triangle, eng.Sphere, eng.Material, eng.Light, None,
               eng.SynthConfig, eng.ProceduralConfig, eng.ServerConfig, eng.ServerStats, eng.ImageMetrics,
               eng.BilateralConfig, eng.AtrousConfig]
    for k, m in enumerate(mirrors):
        want = eng.SAMPLE_RECORD.itemsize if m is None else C.sizeof(m)
        assert out[k] == want, (k, m, out[k], want)


def test_defaults_and_validation_without_device():
    cfg = eng.default_config(image_width=64, image_height=48)
    assert cfg.samples_per_pixel >= 1 and cfg.rpf_tile_size == 8 and cfg.max_depth == 0
    assert eng.job_bytes(cfg) > 64 * 48 * 4
    # any positive tile size is accepted, as by the reference's config check (core/config.h:136-138)
    big = eng.default_config(image_width=64, image_height=48, rpf_tile_size=70)
    assert eng.job_bytes(big) > 0
    bad = eng.default_config(image_width=64, image_height=48, rpf_tile_size=0)
    try:
        eng.job_bytes(bad)
        raise AssertionError("a tile size of 0 was accepted")
    except eng.EngineError as e:
        assert "rpf_tile_size must be positive" in str(e)
