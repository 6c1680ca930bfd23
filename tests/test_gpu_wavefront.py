"""Wavefront trace of flat scenes (rpd_ctx_set_trace_wavefront) against the
path-regeneration kernel and the oracle.

The wavefront form splits each step of a path lane (path_step: resolve the
answered queries, write a finished sample, hand out pixels, start a camera ray
or shade one bounce) from the intersection of its queries, with the lane state
in device memory in between, and hands the slots still running after its last
round to the path-regeneration kernel. Every random draw, hit and record is
formed by the same code in the same order, so the records must equal the
path-regeneration kernel's bit for bit, whatever the number of rounds (a
quarter round per sample leaves most of the pass to the finishing kernel, 64
rounds per sample finish every path inside the wavefront). The reference's
sample loop is main.cpp:649-737 (base pass) and 1451-1538 (adaptive pass)
around shade_mis (main.cpp:394-533).
"""
import ctypes as C

import numpy as np
import pytest

import oracle_bind as ob
import rpd_engine as eng
from parity_util import RECORD_OUTLIER_BUDGET, check_records, engine_camera, engine_config

pytestmark = pytest.mark.gpu

ROUNDS = (0.25, 3.0, 64.0)


@pytest.fixture(scope="module")
def ctxs():
    plain = eng.Context(0)
    wf = eng.Context(0)
    yield plain, wf
    wf.close()
    plain.close()


def _c1_small(w, h, spp=16):
    """The full-size cornell_box settings (16 + 3x4 spp, radius 10, 3 variance
    iterations, aperture 0.14 focused at 1.05) at a small image size."""
    rc = ob.config_c0(w, h)
    rc.spp = spp
    rc.color_window_radius = 10
    rc.var_iterations = 3
    rc.aperture = 0.14
    rc.focus_distance = 1.05
    return rc


def _pass(ctx, rc, seed, scene="cornell_box", counts=None):
    j = eng.Job(ctx, engine_config(rc), engine_camera(rc), seed, scene=scene)
    try:
        j.run_pass(counts, statmc=True)
        return j.get("records"), j.counters()
    finally:
        j.release()


def _same_records(a, b, what):
    assert a.shape == b.shape, (what, a.shape, b.shape)
    diff = (a.view(np.uint32) != b.view(np.uint32)).any(axis=1)
    assert not diff.any(), f"{what}: {int(diff.sum())} of {len(diff)} records differ, first at {int(np.argmax(diff))}"


@pytest.mark.parametrize("rounds", ROUNDS)
def test_base_pass_records_equal_path_regeneration(ctxs, rounds):
    plain, wf = ctxs
    rc = _c1_small(80, 60)
    ref, ref_cnt = _pass(plain, rc, 1234)
    wf.set_trace_wavefront(rounds)
    got, got_cnt = _pass(wf, rc, 1234)
    print(f"rounds {rounds}: {len(got)} samples, {got_cnt['queries']} queries")
    _same_records(ref, got, f"rounds {rounds}")
    assert got_cnt == ref_cnt


def test_uneven_counts_records_equal_path_regeneration(ctxs):
    """An adaptive pass's plan: counts 0..32 per pixel, so some slots run eight
    times as long as the mean and the finishing kernel takes them over."""
    plain, wf = ctxs
    rc = _c1_small(72, 56)
    rng = np.random.default_rng(5)
    counts = rng.integers(0, 5, size=rc.width * rc.height).astype(np.int32)
    counts[rng.random(counts.size) < 0.03] = 32
    ref, ref_cnt = _pass(plain, rc, 99, counts=counts)
    wf.set_trace_wavefront(3.0)
    got, got_cnt = _pass(wf, rc, 99, counts=counts)
    _same_records(ref, got, "uneven counts")
    assert got_cnt == ref_cnt


BUFFERS = ("radiance", "raw_radiance", "denoised", "albedo", "normal", "depth", "variance_denoised")


def test_whole_job_equals_path_regeneration(ctxs):
    plain, wf = ctxs
    rc = _c1_small(64, 64)
    rc.adaptive_passes = 3
    rc.adaptive_spp = 4
    out = []
    for ctx, rounds in ((plain, None), (wf, 3.0)):
        if rounds is not None:
            ctx.set_trace_wavefront(rounds)
        j = eng.Job(ctx, engine_config(rc), engine_camera(rc), 1234, submit=True)
        try:
            j.wait()
            out.append({b: j.get(b) for b in BUFFERS} | {"counters": j.counters()})
        finally:
            j.release()
    for b in BUFFERS:
        np.testing.assert_array_equal(out[0][b].view(np.uint32), out[1][b].view(np.uint32), err_msg=b)
    assert out[0]["counters"] == out[1]["counters"]


# name: (camera position, direction, fov, focus distance) from path_tracer/config/<name>.yaml
CAMERAS = {
    "cornell_spheres": ((0.278, 0.2744, -0.8), (0.0, 0.0, 1.0), 38.6, 0.8),
    "veach_mis": ((0.0, 2.0, -8.0), (0.0, -0.15, 1.0), 45.0, 8.0),
    "material_showcase": ((0.0, 3.5, -10.0), (0.0, -0.2, 1.0), 40.0, 10.0),
    "caustics": ((0.0, 3.5, -7.0), (0.0, -0.25, 1.0), 50.0, 7.5),
}


def _scene_config(name, max_depth):
    pos, dirn, fov, focus = CAMERAS[name]
    rc = ob.config_c0(48, 36)
    rc.spp = 6
    rc.cam_pos, rc.cam_dir, rc.fov, rc.focus_distance = pos, dirn, fov, focus
    rc.max_depth = max_depth
    return rc


@pytest.mark.parametrize("max_depth", (0, 8))
@pytest.mark.parametrize("name", sorted(CAMERAS))
def test_other_flat_scenes_equal_path_regeneration(ctxs, name, max_depth):
    """Spheres in the flat loop, mirror, GGX and Beckmann materials, point and
    directional lights, several lights, with and without a depth limit."""
    plain, wf = ctxs
    rc = _scene_config(name, max_depth)
    ref, ref_cnt = _pass(plain, rc, 4321, scene=name)
    wf.set_trace_wavefront(3.0)
    got, got_cnt = _pass(wf, rc, 4321, scene=name)
    _same_records(ref, got, f"{name} max_depth {max_depth}")
    assert got_cnt == ref_cnt


def _env_scene(ctx, key, env):
    """material_showcase with a constant environment (the environment shadow
    ray and the environment MIS weights), created through rpd_scene_create."""
    if key not in ctx._scenes:
        tris, sph, mats, lights, _, n = eng.builtin_scene_arrays("material_showcase")
        e = (C.c_float * 3)(*env)
        s = C.c_void_p()
        eng.check(ctx.L.rpd_scene_create(ctx.h, tris, n[0], sph, n[1], mats, n[2], lights, n[3], e, C.byref(s)))
        ctx._scenes[key] = s
    return key


def test_environment_lit_scene_equals_path_regeneration_and_oracle(ctxs):
    plain, wf = ctxs
    env = (0.2, 0.25, 0.3)
    rc = _scene_config("material_showcase", 8)
    ref, ref_cnt = _pass(plain, rc, 77, scene=_env_scene(plain, "showcase_env", env))
    wf.set_trace_wavefront(3.0)
    got, got_cnt = _pass(wf, rc, 77, scene=_env_scene(wf, "showcase_env", env))
    _same_records(ref, got, "environment-lit material_showcase")
    assert got_cnt == ref_cnt
    tris, sph, mats, lights, _, n = eng.builtin_scene_arrays("material_showcase")
    arrays = (tris, sph, mats, lights, (C.c_float * 3)(*env), n)
    oj = ob.OracleJob(rc, 77, nthreads=1, keep_records=True, scene=arrays)
    oj.run_pass(np.full(rc.width * rc.height, rc.spp, np.int32), statmc=True)
    check_records(oj.get("records"), got, RECORD_OUTLIER_BUDGET["procedural"], "environment-lit:")


def test_cornell_records_match_oracle(ctxs):
    """The wavefront records against the oracle directly, at the full-size
    settings: draws, flags and light index exact, radiance within 1e-4,
    rounding-boundary outliers within the cornell_box budget."""
    _, wf = ctxs
    rc = _c1_small(96, 72)
    wf.set_trace_wavefront(3.0)
    got, _ = _pass(wf, rc, 1234)
    oj = ob.OracleJob(rc, 1234, nthreads=1, keep_records=True)
    oj.run_pass(np.full(rc.width * rc.height, rc.spp, np.int32), statmc=True)
    check_records(oj.get("records"), got, RECORD_OUTLIER_BUDGET["cornell_box"], "wavefront cornell:")


def test_rounds_argument_is_checked(ctxs):
    _, wf = ctxs
    for bad in (-1.0, 65.0, float("nan")):
        with pytest.raises(eng.EngineError):
            wf.set_trace_wavefront(bad)
    wf.set_trace_wavefront(0.0)
