"""Gather overlap: a pass's RPF grid gather runs on a second stream beside the
next pass's trace, the passes alternating between two plan and record sets.

The grid's FP64 sums must not depend on it: every gather keeps its pass order
and reads only its own pass's plan and records. Jobs run with the overlap on
and off are compared bit for bit over every buffer the pipeline writes.
"""
import ctypes as C

import numpy as np
import pytest

import oracle_bind as ob
import rpd_engine as eng
from parity_util import engine_camera, engine_config

pytestmark = pytest.mark.gpu

BUFFERS = ("grid", "tiles", "stats", "sums", "radiance", "raw_radiance", "denoised", "variance_denoised",
           "uncertainty", "alpha", "sample_count", "hit_count", "pass_counts", "records")


def _bits(a):
    a = np.ascontiguousarray(a)
    return a.view(np.uint64 if a.dtype == np.float64 else np.uint32)


def _run(ctx, rc, seed, overlap, scene="cornell_box"):
    ctx.set_gather_overlap(overlap)
    j = eng.Job(ctx, engine_config(rc), engine_camera(rc), seed, scene=scene, submit=True)
    try:
        j.wait()
        return {n: _bits(j.get(n)) for n in BUFFERS}
    finally:
        j.release()


def _same(a, b, what):
    for n in BUFFERS:
        assert a[n].shape == b[n].shape, f"{what}: {n} shape"
        bad = np.flatnonzero(a[n].ravel() != b[n].ravel())
        assert bad.size == 0, f"{what}: {n} differs at {bad.size} words, first {bad[:8]}"


@pytest.mark.parametrize("w,h,passes,spp,ap", [(128, 96, 3, 16, 0.14), (80, 64, 1, 4, 0.0), (72, 56, 4, 8, 0.14)])
def test_cornell_jobs_same_bits_with_and_without_overlap(w, h, passes, spp, ap):
    """The C1 settings (16 base spp, 3 adaptive passes of 4, StatMC r 10, 3
    variance iterations, aperture 0.14), a single adaptive pass (only the base
    pass's gather overlaps) and four passes (both sets reused twice)."""
    rc = ob.config_c1(w, h)
    rc.spp, rc.adaptive_passes, rc.aperture = spp, passes, ap
    rc.focus_distance = 1.05 if ap > 0 else 1.0
    ctx = eng.Context(0)
    try:
        off = _run(ctx, rc, 4242, False)
        on = _run(ctx, rc, 4242, True)
        _same(off, on, f"cornell {w}x{h} passes {passes}")
    finally:
        ctx.close()


def test_environment_lit_flat_scene_same_bits_with_and_without_overlap():
    """material_showcase under a constant environment, through the lens, max
    depth 8: every RPF accumulator receives samples (environment and lens
    dimensions included), spheres in the flat loop, GGX, Beckmann and mirror
    materials, several lights."""
    rc = ob.config_c0(72, 48)
    rc.cam_pos, rc.cam_dir, rc.fov, rc.focus_distance = (0.0, 3.5, -10.0), (0.0, -0.2, 1.0), 40.0, 10.0
    rc.spp, rc.adaptive_passes, rc.aperture, rc.max_depth = 8, 3, 0.1, 8
    ctx = eng.Context(0)
    try:
        tris, sph, mats, lights, _, n = eng.builtin_scene_arrays("material_showcase")
        env = (C.c_float * 3)(0.2, 0.25, 0.3)
        s = C.c_void_p()
        eng.check(ctx.L.rpd_scene_create(ctx.h, tris, n[0], sph, n[1], mats, n[2], lights, n[3], env, C.byref(s)))
        ctx._scenes["showcase_env"] = s
        off = _run(ctx, rc, 77, False, scene="showcase_env")
        on = _run(ctx, rc, 77, True, scene="showcase_env")
        _same(off, on, "environment-lit material_showcase 72x48")
    finally:
        ctx.close()


def test_bvh_scene_keeps_its_gathers_in_line():
    """Scenes traced with a BVH do not take the overlap (the gather slowed their
    trace more than it hid); the setting leaves their results unchanged."""
    pc = eng.procedural_config(seed=5, subdivision_level=3, n_spheres=24)
    rc = ob.RenderConfig(width=80, height=56, spp=8, fov=40.0, cam_pos=(0.0, 0.6, 3.4), cam_dir=(0.0, -0.17, -1.0),
                         focus_distance=3.4, aperture=0.06, adaptive_spp=8, adaptive_passes=4)
    rc.max_depth = 8
    ctx = eng.Context(0)
    try:
        ctx.procedural_scene("procedural-overlap", pc)
        off = _run(ctx, rc, 99, False, scene="procedural-overlap")
        on = _run(ctx, rc, 99, True, scene="procedural-overlap")
        _same(off, on, "procedural 80x56")
    finally:
        ctx.close()


def test_passes_after_an_overlapped_job_use_the_first_set():
    """A job whose whole pipeline ran with the overlap ends on the second plan
    and record set; a pass run after it through the step interface goes back to
    the first set (whose records hold the base pass) and matches the same pass
    on a job that never overlapped."""
    rc = ob.config_c1(64, 48)
    res = []
    ctx = eng.Context(0)
    try:
        for overlap in (False, True):
            ctx.set_gather_overlap(overlap)
            j = eng.Job(ctx, engine_config(rc), engine_camera(rc), 31, submit=True)
            try:
                j.wait()
                j.run_pass(None, statmc=True)
                res.append({n: _bits(j.get(n)) for n in ("records", "pass_counts", "grid", "stats")})
            finally:
                j.release()
    finally:
        ctx.close()
    for n in res[0]:
        np.testing.assert_array_equal(res[0][n], res[1][n], err_msg=n)


def test_served_jobs_with_overlap_equal_jobs_alone():
    """Jobs admitted together by the request loop (kernel by kernel, so every
    job overlaps its gathers) against the same jobs run alone without it."""
    reqs = []
    for k, (w, h, spp, passes, ap) in enumerate([(64, 48, 4, 3, 0.0), (96, 64, 8, 2, 0.14), (80, 56, 4, 4, 0.14)]):
        rc = ob.config_c0(w, h)
        rc.spp, rc.adaptive_passes, rc.aperture, rc.focus_distance = spp, passes, ap, 1.05
        reqs.append(rc)
    ctx = eng.Context(0)
    try:
        alone = [_run(ctx, rc, 500 + k, False) for k, rc in enumerate(reqs)]
        ctx.set_gather_overlap(True)
        srv = eng.Server(ctx, max_in_flight=3, graph_passes=False)
        try:
            ids = [srv.submit(engine_config(rc), engine_camera(rc), 500 + k) for k, rc in enumerate(reqs)]
            for k, rid in enumerate(ids):
                j, _ = srv.wait(rid)
                try:
                    got = {n: _bits(j.get(n)) for n in BUFFERS}
                finally:
                    j.release()
                _same(alone[k], got, f"served request {k}")
        finally:
            srv.close()
    finally:
        ctx.close()
