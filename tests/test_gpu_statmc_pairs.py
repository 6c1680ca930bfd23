"""StatMC window as two passes: a pair pass computes each pair of the window
once (a pair's weight is the same bits with its two pixels exchanged) into a
table in the job's tape buffer, and a sum pass adds every pixel's full window
in the reference's order (main.cpp:1032-1126).

Jobs run with the two passes and with the one-pass window are compared bit for
bit over every buffer StatMC writes, for several radii (the largest the tape
buffer holds, and one beyond, which keeps one pass), every tile shape, the
debug outputs, and a flat scene with an environment and a lens.
"""
import ctypes as C

import numpy as np
import pytest

import oracle_bind as ob
import rpd_engine as eng
from parity_util import engine_camera, engine_config

pytestmark = pytest.mark.gpu

BUFFERS = ("denoised", "uncertainty", "alpha", "var_total", "var_eff", "var_ratio", "variance_denoised", "radiance",
           "sample_count")
DEBUG = ("debug_neighbors", "debug_weight_sum", "debug_neff")


def _bits(a):
    a = np.ascontiguousarray(a)
    return a.view(np.uint64 if a.dtype == np.float64 else np.uint32)


def _run(ctx, rc, seed, pairs, scene="cornell_box", names=BUFFERS):
    ctx.set_statmc_pairs(pairs)
    j = eng.Job(ctx, engine_config(rc), engine_camera(rc), seed, scene=scene, submit=True)
    try:
        j.wait()
        return {n: _bits(j.get(n)) for n in names}
    finally:
        j.release()


def _same(a, b, what):
    for n in a:
        assert a[n].shape == b[n].shape, f"{what}: {n} shape"
        bad = np.flatnonzero(a[n].ravel() != b[n].ravel())
        assert bad.size == 0, f"{what}: {n} differs at {bad.size} words, first {bad[:8]}"


@pytest.mark.parametrize("radius", [1, 3, 10, 13, 14])
def test_pair_passes_same_bits_as_one_pass(radius):
    """The C1 settings at 96x72 (16 base spp, 3 adaptive passes of 4, 3
    variance iterations, aperture 0.14) with the window radius varied; 13 is
    the largest whose pair table fits the tape buffer, 14 keeps one pass."""
    rc = ob.config_c1(96, 72)
    rc.color_window_radius = radius
    ctx = eng.Context(0)
    try:
        one = _run(ctx, rc, 1234, False)
        two = _run(ctx, rc, 1234, True)
        _same(one, two, f"radius {radius}")
    finally:
        ctx.close()


def test_every_tile_shape_and_the_debug_outputs():
    rc = ob.config_c1(80, 56)
    names = BUFFERS + DEBUG
    ctx = eng.Context(0)
    try:
        ctx.set_debug_outputs(True)
        ref = _run(ctx, rc, 77, False, names=names)
        for tw, th in eng.statmc_tile_shapes():
            ctx.set_statmc_tile(tw, th)
            _same(ref, _run(ctx, rc, 77, True, names=names), f"shape {tw}x{th}")
    finally:
        ctx.close()


def test_environment_lit_flat_scene():
    """material_showcase under a constant environment, through the lens, max
    depth 8, radius 7: other materials and reliable components than cornell's."""
    rc = ob.config_c0(72, 48)
    rc.cam_pos, rc.cam_dir, rc.fov, rc.focus_distance = (0.0, 3.5, -10.0), (0.0, -0.2, 1.0), 40.0, 10.0
    rc.spp, rc.adaptive_passes, rc.aperture, rc.max_depth, rc.color_window_radius = 8, 2, 0.1, 8, 7
    ctx = eng.Context(0)
    try:
        tris, sph, mats, lights, _, n = eng.builtin_scene_arrays("material_showcase")
        env = (C.c_float * 3)(0.2, 0.25, 0.3)
        s = C.c_void_p()
        eng.check(ctx.L.rpd_scene_create(ctx.h, tris, n[0], sph, n[1], mats, n[2], lights, n[3], env, C.byref(s)))
        ctx._scenes["showcase_env"] = s
        one = _run(ctx, rc, 5, False, scene="showcase_env")
        two = _run(ctx, rc, 5, True, scene="showcase_env")
        _same(one, two, "environment-lit material_showcase")
    finally:
        ctx.close()
