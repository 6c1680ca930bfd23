"""The procedural scene on the GPU (BVH traversal, GGX spheres, four area
lights, environment, thin lens, max depth 8) against the oracle tracing the
same scene arrays: per-sample records of identical random streams.
"""
import numpy as np
import pytest

import oracle_bind as ob
import rpd_engine as eng
from parity_util import RECORD_OUTLIER_BUDGET, check_records, engine_camera, engine_config, rel_err

pytestmark = pytest.mark.gpu


def procedural_render_config(w, h, spp):
    rc = ob.RenderConfig(width=w, height=h, spp=spp, fov=40.0, cam_pos=(0.0, 0.6, 3.4),
                         cam_dir=(0.0, -0.17, -1.0), focus_distance=3.4, aperture=0.06,
                         adaptive_spp=4, adaptive_passes=1)
    rc.max_depth = 8
    return rc


@pytest.mark.parametrize("level,spheres,seed", [(3, 40, 7), (4, 24, 11)])
def test_procedural_scene_records_match_oracle(engine_ctx, level, spheres, seed):
    pc = eng.procedural_config(seed=seed, subdivision_level=level, n_spheres=spheres)
    arrays = eng.procedural_arrays(pc)
    rc = procedural_render_config(40, 30, 3)
    oj = ob.OracleJob(rc, 1234, nthreads=1, keep_records=True, scene=arrays)
    counts = np.full(rc.width * rc.height, rc.spp, np.int32)
    oj.run_pass(counts, statmc=True)
    orec = oj.get("records")
    name = f"procedural-{seed}-{level}-{spheres}"
    engine_ctx.procedural_scene(name, pc)
    gj = eng.Job(engine_ctx, engine_config(rc), engine_camera(rc), 1234, scene=name)
    gj.run_pass(None, statmc=True)
    grec = gj.get("records")
    name_ = f"procedural-{seed}-{level}-{spheres}:"
    check_records(orec, grec, RECORD_OUTLIER_BUDGET["procedural"], name_)
    print("env captured", int((orec[:, 28].astype(int) & 32 != 0).sum()), "hits", int(orec[:, 5].sum()))
    assert (orec[:, 28].astype(int) & 32 != 0).any()      # the environment dimension is exercised


def _rays(rng, n, spread=4.0):
    o = rng.uniform(-spread, spread, (n, 3)).astype(np.float32)
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    aim = rng.random(n) < 0.25
    tgt = rng.uniform(-0.5, 0.5, (n, 3))
    d[aim] = (tgt[aim] - o[aim]) / np.linalg.norm(tgt[aim] - o[aim], axis=1, keepdims=True)
    ax = rng.random(n) < 0.05
    d[ax] = 0.0
    d[ax, rng.integers(0, 3, ax.sum())] = rng.choice([-1.0, 1.0], ax.sum())
    return o, d.astype(np.float32)


@pytest.mark.parametrize("level,spheres", [(3, 40), (5, 64), (6, 8)])
def test_device_built_bvh_gives_the_exhaustive_hits(level, spheres):
    """The LBVH built on the device (downloaded with the scene for host queries)
    against the exhaustive loop over every primitive: identical hit, distance,
    primitive and sphere flag."""
    ctx = eng.Context(0)
    try:
        ctx.set_bvh_builder(eng.RPD_BVH_DEVICE_LBVH)
        name = f"lbvh-{level}-{spheres}"
        ctx.procedural_scene(name, eng.procedural_config(subdivision_level=level, n_spheres=spheres))
        assert ctx.bvh_builder(name) == eng.RPD_BVH_DEVICE_LBVH
        nodes, depth = ctx.bvh_info(name)
        assert nodes > 0 and 1 <= depth <= 16
        rng = np.random.default_rng(100 + level)
        n = 20000
        o, d = _rays(rng, n)
        tmax = np.where(rng.random(n) < 0.5, np.inf, rng.uniform(0.1, 6.0, n)).astype(np.float32)
        occ = (rng.random(n) < 0.4).astype(np.int32)
        a = ctx.scene_query(name, o, d, tmax, occ, use_bvh=True)
        b = ctx.scene_query(name, o, d, tmax, occ, use_bvh=False)
        np.testing.assert_array_equal(a[0], b[0], err_msg="hit")
        cl = occ == 0
        for x, y, what in zip(a[1:], b[1:], ("t", "index", "is_sphere")):
            np.testing.assert_array_equal(x[cl], y[cl], err_msg=what)
        assert a[0].mean() > 0.2
    finally:
        ctx.close()


def test_device_built_bvh_renders_the_same_records():
    """Hits do not depend on the tree, so a pass on the device-built BVH writes
    the same records as on the host-built one, bit for bit."""
    pc = eng.procedural_config(seed=11, subdivision_level=4, n_spheres=24)
    rc = procedural_render_config(40, 30, 3)
    recs = []
    for builder in (eng.RPD_BVH_HOST_SAH, eng.RPD_BVH_DEVICE_LBVH):
        ctx = eng.Context(0)
        try:
            ctx.set_bvh_builder(builder)
            ctx.procedural_scene("p", pc)
            assert ctx.bvh_builder("p") == builder
            gj = eng.Job(ctx, engine_config(rc), engine_camera(rc), 1234, scene="p")
            gj.run_pass(None, statmc=True)
            recs.append(gj.get("records"))
            gj.release()
        finally:
            ctx.close()
    np.testing.assert_array_equal(recs[0], recs[1])
