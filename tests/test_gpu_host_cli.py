"""The path_tracer executable end to end on the GPU: the reference's command
line (path_tracer/main.cpp:1566-1922) renders a small cornell_box job.

Checked:
  * exit status 0 and the stage log lines scripts/benchmark_pass_timing.sh
    parses (main.cpp:1543-1564 stages);
  * the output file set of output_buffers (core/buffers.h:197-250) in
    <dir>/<leaf>*, and the debug_statmc_outputs images in the working
    directory (main.cpp:1129-1157, 1283-1289, 1319-1321);
  * every .hdr is a Radiance RGBE file (decoded here from the format's
    definition, flat and run-length scanlines) that matches the engine's float
    buffer of the same job within RGBE precision (8-bit mantissa of a shared
    exponent: error below 2^-7 of the pixel's largest channel);
  * the PNG files are 8-bit RGB images of the right size with valid chunk CRCs
    and image data that inflates to the expected scanline bytes.
"""
import os
import re
import struct
import subprocess
import tempfile
import zlib

import numpy as np
import pytest

import rpd_engine as eng

pytestmark = pytest.mark.gpu

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(REPO, "rpd-statmc-denoiser_amd", "bin", "path_tracer")
W, H, SEED = 48, 40, 1234

CONFIG = f"""image_width: {W}
image_height: {H}
samples_per_pixel: 4
scene: cornell_box
output: cornell
fov: 38.6
camera_position: [0.278, 0.2744, -0.8]
camera_direction: [0.0, 0.0, 1.0]
focus_distance: 0.8
statmc_enabled: true
color_window_radius: 5
adaptive_spp: 4
adaptive_base_samples: 1
adaptive_passes: 2
debug_statmc_outputs: true
"""

MAIN_FILES = [".png", ".hdr", "_raw.png", "_raw.hdr", "_sensitivity.hdr", "_sensitivity_pixel.hdr",
              "_sensitivity_brdf.hdr", "_sensitivity_lens.hdr", "_sensitivity_light.hdr",
              "_sensitivity_light_uv.hdr", "_sensitivity_light_select.hdr", "_sensitivity_environment.hdr",
              "_sensitivity_rr.hdr", "_sensitivity_confidence.hdr", "_sensitivity_gradient.hdr", "_uncertainty.hdr",
              "_alpha.hdr", "_vardenoised.hdr", "_adaptive_importance.hdr", "_sample_fraction.hdr", "_albedo.hdr",
              "_normal.hdr", "_worldpos.hdr", "_depth.hdr", "_var_total.hdr", "_var_eff.hdr", "_var_ratio.hdr"]
DEBUG_FILES = ["neighbors", "alpha", "var_total", "var_eff", "var_ratio", "weight_sum", "neff", "sensitivity_all",
               "sensitivity_pixel", "sensitivity_brdf", "sensitivity_lens", "sensitivity_light",
               "sensitivity_light_uv", "sensitivity_light_select", "sensitivity_environment", "sensitivity_rr",
               "sensitivity_confidence", "sensitivity_gradient", "reliability_base", "reliability_denoise",
               "adaptive_extra_counts", "variance_source_adaptive"]
LOG_LINES = ["Rendering completed in", "Adaptive sample count computation completed in",
             "Adaptive rendering completed in", "Sensitivity recompute completed in", "StatMC denoise completed in",
             "Sample-variance denoise completed in"]


def read_rgbe(path):
    """Radiance .hdr reader: header lines, the -Y H +X W resolution line, then
    scanlines either flat (4 bytes per pixel) or new-style run-length encoded
    (2 2 hi lo marker, then each of the 4 components in runs: count > 128 is a
    run of count - 128 equal bytes, otherwise count literal bytes)."""
    data = open(path, "rb").read()
    assert data.startswith(b"#?RADIANCE") or data.startswith(b"#?RGBE")
    pos = 0
    fmt_ok = False
    while True:
        end = data.index(b"\n", pos)
        line = data[pos:end]
        pos = end + 1
        if line.startswith(b"FORMAT="):
            fmt_ok = line == b"FORMAT=32-bit_rle_rgbe"
        if line == b"":
            break
    assert fmt_ok
    end = data.index(b"\n", pos)
    m = re.match(rb"-Y (\d+) \+X (\d+)", data[pos:end])
    h, w = int(m.group(1)), int(m.group(2))
    pos = end + 1
    out = np.zeros((h, w, 4), np.uint8)
    for y in range(h):
        if 8 <= w < 32768 and data[pos] == 2 and data[pos + 1] == 2 and (data[pos + 2] << 8 | data[pos + 3]) == w:
            pos += 4
            for c in range(4):
                x = 0
                while x < w:
                    cnt = data[pos]
                    pos += 1
                    if cnt > 128:
                        out[y, x:x + cnt - 128, c] = data[pos]
                        pos += 1
                        x += cnt - 128
                    else:
                        out[y, x:x + cnt, c] = np.frombuffer(data[pos:pos + cnt], np.uint8)
                        pos += cnt
                        x += cnt
        else:
            out[y] = np.frombuffer(data[pos:pos + 4 * w], np.uint8).reshape(w, 4)
            pos += 4 * w
    assert pos == len(data)
    e = out[:, :, 3].astype(np.int32)
    scale = np.where(e > 0, np.ldexp(1.0, e - (128 + 8)), 0.0)
    return (out[:, :, :3].astype(np.float64) + 0.5 * (e[:, :, None] > 0)) * scale[:, :, None], (h, w)


def check_png(path, w, h):
    data = open(path, "rb").read()
    assert data[:8] == b"\x89PNG\r\n\x1a\n"
    pos, idat, ihdr = 8, b"", None
    while pos < len(data):
        n, kind = struct.unpack(">I4s", data[pos:pos + 8])
        body = data[pos + 8:pos + 8 + n]
        crc, = struct.unpack(">I", data[pos + 8 + n:pos + 12 + n])
        assert zlib.crc32(kind + body) & 0xffffffff == crc, kind
        if kind == b"IHDR":
            ihdr = struct.unpack(">IIBBBBB", body)
        elif kind == b"IDAT":
            idat += body
        pos += 12 + n
    assert ihdr[:4] == (w, h, 8, 2)                 # 8-bit truecolour
    assert len(zlib.decompress(idat)) == h * (1 + 3 * w)


def _engine_job(ctx):
    cfg = eng.default_config(image_width=W, image_height=H, samples_per_pixel=4, use_statmc=1, color_window_radius=5,
                             adaptive_spp=4, adaptive_base_samples=1, adaptive_passes=2)
    cam = eng.make_camera((0.278, 0.2744, -0.8), (0.0, 0.0, 1.0), (0.0, 1.0, 0.0), 38.6, 0.8, 0.0)
    j = eng.Job(ctx, cfg, cam, SEED, submit=True)
    j.wait()
    return j


def _close(decoded, ref, what):
    ref = np.asarray(ref, np.float64).reshape(decoded.shape[0], decoded.shape[1], -1)
    if ref.shape[2] == 1:
        ref = np.repeat(ref, 3, axis=2)
    mx = np.maximum(np.abs(ref).max(axis=2, keepdims=True), 1e-30)
    small = mx < 1e-31                              # encoded as zero
    err = np.abs(decoded - np.maximum(ref, 0.0))
    ok = (err <= mx * 2.0 ** -7) | small
    assert ok.all(), (what, float((err / mx).max()))


def test_path_tracer_renders_the_reference_file_set():
    with tempfile.TemporaryDirectory() as tmp:
        cfg_path = os.path.join(tmp, "small.yaml")
        open(cfg_path, "w").write(CONFIG)
        out_dir = os.path.join(tmp, "renders", "cornell_small")
        print("running path_tracer", flush=True)
        r = subprocess.run([BIN, "-c", cfg_path, "-o", out_dir, "--seed", str(SEED)], cwd=tmp, capture_output=True,
                           text=True, timeout=120)
        log = r.stdout + r.stderr
        assert r.returncode == 0, log[-2000:]
        for line in LOG_LINES:
            assert line in log, line
        assert "Successfully saved output to" in log
        base = os.path.join(out_dir, "cornell_small")
        missing = [s for s in MAIN_FILES if not os.path.exists(base + s)]
        assert not missing, missing
        missing = [f for f in DEBUG_FILES if not os.path.exists(os.path.join(tmp, f"debug_{f}.hdr"))]
        assert not missing, missing
        for png in (".png", "_raw.png"):
            check_png(base + png, W, H)
        # the same job through the library: every .hdr is its float buffer within RGBE precision
        ctx = eng.Context(0)
        try:
            ctx.set_debug_outputs(True)
            j = _engine_job(ctx)
            pairs = {".hdr": "radiance", "_raw.hdr": "raw_radiance", "_albedo.hdr": "albedo",
                     "_normal.hdr": "normal", "_worldpos.hdr": "world_pos", "_depth.hdr": "depth",
                     "_uncertainty.hdr": "uncertainty", "_alpha.hdr": "alpha", "_vardenoised.hdr": "variance_denoised",
                     "_sensitivity.hdr": "sens_all", "_sensitivity_gradient.hdr": "sensitivity_gradient",
                     "_sample_fraction.hdr": "sample_fraction", "_var_total.hdr": "var_total"}
            for suffix, name in pairs.items():
                img, (h, w) = read_rgbe(base + suffix)
                assert (h, w) == (H, W)
                _close(img, j.get(name), suffix)
            for f, name in [("neighbors", "debug_neighbors"), ("weight_sum", "debug_weight_sum"),
                            ("neff", "debug_neff"), ("variance_source_adaptive", "debug_variance_source")]:
                img, _ = read_rgbe(os.path.join(tmp, f"debug_{f}.hdr"))
                _close(img, j.get(name), f)
            counts, _ = read_rgbe(os.path.join(tmp, "debug_adaptive_extra_counts.hdr"))
            _close(counts, j.get("pass_counts").astype(np.float32), "extra counts")
            nb = j.get("debug_neighbors")
            assert nb.max() >= 2 and (nb == np.floor(nb)).all()
            j.release()
        finally:
            ctx.close()
