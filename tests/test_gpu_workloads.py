"""The benchmark workloads end to end on the GPU against the oracle.

Two kinds of test per workload:

* lock-step: both sides run the same stage sequence as render()
  (path_tracer/main.cpp:1543-1564), and before every stage the engine is handed
  the oracle's state, so each stage is compared on identical inputs:
  - every sampling pass: per-sample records of identical RNG streams (draws,
    flags and light index exact, radiance within 1e-4 relative), the pixel
    statistics and the RPF grid;
  - variance smoothing: bit-exact;
  - adaptive allocation: counts bit-exact (the importance field within one
    float ulp: the smoothing's expf is correctly rounded here, glibc's is not
    on about 1e-4 of arguments);
  - sensitivity tiles within 1e-5, StatMC with no pixel beyond 1e-4.
* free-running: one whole job (rpd_job_submit, or the serving loop) against
  the oracle's render(), compared on the final buffers.

Outliers: a sample whose path takes the other branch at a rounding boundary
(an intersection within an ulp of an edge, a Russian-roulette or light-CDF
draw within an ulp of the threshold) legitimately differs. Such samples are
counted, reported and held to a budget (parity_util.RECORD_OUTLIER_BUDGET). In a free-running job a
differing sample changes its pixel's statistics and, through the importance
field, changes the next pass's allocation. The allocation is one sequential
scan over the whole image (main.cpp:1373-1438): a different importance in one
pixel moves the running sums, and with them the systematic residual draw of
every later pixel. A free-running job whose earlier pass held one outlier
therefore samples many pixels differently from the oracle from then on, as two
runs of the reference would after any perturbation. Free-running jobs are
compared exactly when no pass diverged, and otherwise statistically: the total
sample count is equal, and the GPU image differs from the oracle's by much less
than two oracle runs with different seeds differ (relMSE ratio below
parity_util.DIVERGED_RELMSE_RATIO), with an image mean within 1e-3.
"""
import numpy as np
import pytest

import oracle_bind as ob
import rpd_engine as eng
from parity_util import (RECORD_OUTLIER_BUDGET, compare_whole_job, engine_camera, engine_config, records_outliers,
                         rel_err)

pytestmark = pytest.mark.gpu

SEED = 1234


def _inject_state(oj, gj):
    """Hands the engine the oracle's per-pixel state and RPF grid."""
    gj.set("stats", oj.get("stats"))
    sums = np.concatenate([oj.get("radiance"), oj.get("albedo"), oj.get("normal"), oj.get("world_pos"),
                           oj.get("depth")[:, None]], axis=1)
    gj.set("sums", sums)
    gj.set("hit_count", oj.get("hit_count"))
    gj.set("sample_count", oj.get("sample_count"))
    gj.set("grid", oj.get("grid"))


def _check_pass(oj, gj, what, log, budget):
    orec, grec = oj.get("records"), gj.get("records")
    bad = records_outliers(orec, grec)
    n_bad = int(bad.sum())
    log.append(f"{what}: samples {len(orec)} outliers {n_bad} (budget {budget * len(orec):.1f})")
    assert n_bad <= budget * len(orec), log[-1]
    assert rel_err(orec[~bad, 1:4], grec[~bad, 1:4]).max() <= 1e-4
    # pixels holding an outlier sample may differ; every other pixel's statistics match
    out_px = np.zeros(oj.w * oj.ht, bool)
    out_px[orec[bad, 0].astype(np.int64)] = True
    np.testing.assert_array_equal(oj.get("sample_count"), gj.get("sample_count"))
    os_, gs = oj.get("stats"), gj.get("stats")
    np.testing.assert_array_equal(os_[~out_px, 2].view(np.int32), gs[~out_px, 2].view(np.int32))
    cols = [0, 3, 4, 5, 6, 7, 8]
    px_bad = (rel_err(os_[:, cols], gs[:, cols], floor=1e-3) > 1e-4).any(axis=1)
    assert not (px_bad & ~out_px).any(), f"{what}: {int((px_bad & ~out_px).sum())} pixel statistics differ"
    # the grid: its accumulators are sums of non-negative per-sample terms, so
    # each is within the per-sample radiance tolerance (1e-4 relative) unless
    # an outlier sample splatted into its node (at most 4 nodes per sample)
    og, gg = oj.get("grid"), gj.get("grid")
    assert og.shape == gg.shape
    node_err = (np.abs(og - gg) / np.maximum(np.abs(og), 1e-300)).max(axis=1)
    node_err[np.abs(og - gg).max(axis=1) == 0.0] = 0.0
    n_nodes_bad = int((node_err > 1e-4).sum())
    log.append(f"{what}: grid max rel err {node_err.max():.2e}, nodes beyond 1e-4 {n_nodes_bad}")
    assert n_nodes_bad <= 4 * n_bad, log[-1]


def _lockstep(ctx, rc, seed, scene="cornell_box", oracle_scene=None, nthreads=8, budget="cornell_box"):
    n = rc.width * rc.height
    oj = ob.OracleJob(rc, seed, nthreads=nthreads, keep_records=True, scene=oracle_scene)
    gj = eng.Job(ctx, engine_config(rc), engine_camera(rc), seed, scene=scene)
    log = []
    try:
        oj.run_pass(np.full(n, rc.spp, np.int32), statmc=True)
        gj.run_pass(None, statmc=True)
        _check_pass(oj, gj, "base pass", log, RECORD_OUTLIER_BUDGET[budget])
        for p in range(rc.adaptive_passes):
            _inject_state(oj, gj)
            oj.variance()
            gj.variance()
            np.testing.assert_array_equal(oj.get("variance_denoised"), gj.get("variance_denoised"))
            oc = oj.adaptive_counts()
            gc = gj.adaptive_counts()
            oi, gi = oj.get("adaptive_importance"), gj.get("adaptive_importance")
            ulps = np.abs(oi.view(np.int32).astype(np.int64) - gi.view(np.int32).astype(np.int64))
            assert ulps.max() <= 1
            np.testing.assert_array_equal(oc, gc, err_msg=f"adaptive pass {p} counts")
            assert oc.sum() <= rc.adaptive_spp * n
            log.append(f"pass {p + 1}: {int(oc.sum())} samples allocated, importance ulps>0 {int((ulps > 0).sum())}")
            oj.run_pass(oc, statmc=True)
            gj.run_pass(oc, statmc=True)
            _check_pass(oj, gj, f"adaptive pass {p + 1}", log, RECORD_OUTLIER_BUDGET[budget])
        _inject_state(oj, gj)
        oj.recompute_sensitivity()
        gj.recompute_sensitivity()
        terr = np.abs(oj.get("tiles") - gj.get("tiles")).max()
        log.append(f"sensitivity: tile max abs err {terr:.2e}")
        assert terr <= 1e-5
        np.testing.assert_allclose(gj.get("sensitivity_gradient"), oj.get("sensitivity_gradient"), atol=1e-5)
        np.testing.assert_allclose(gj.get("sensitivity_confidence"), oj.get("sensitivity_confidence"), atol=1e-5)
        gj.set("tiles", oj.get("tiles").ravel())
        gj.set("sensitivity_gradient", oj.get("sensitivity_gradient"))
        oj.statmc()
        gj.statmc()
        od, gd = oj.get("denoised"), gj.get("denoised")
        e = rel_err(od, gd).max(axis=1)
        exact = (od == gd).all(axis=1).mean()
        log.append(f"statmc: max rel err {e.max():.2e}, bit-identical pixels {exact:.4f}")
        assert (e > 1e-4).sum() == 0, log[-1]
        np.testing.assert_allclose(gj.get("alpha"), oj.get("alpha"), rtol=0, atol=1e-6)
        assert rel_err(oj.get("uncertainty"), gj.get("uncertainty"), floor=1e-12).max() <= 1e-5
        oj.variance()
        gj.variance()
        np.testing.assert_array_equal(oj.get("variance_denoised"), gj.get("variance_denoised"))
        oj.finalize()
        gj.finalize()
        # FrameBuffers::finalize (core/buffers.h:152-167), then render() keeps the
        # mean as raw_radiance and outputs the denoised image (main.cpp:1557-1559)
        pairs = [("radiance", "raw_radiance"), ("denoised", "radiance"), ("albedo", "albedo"),
                 ("normal", "normal"), ("depth", "depth")]
        for oname, gname in pairs:
            e = rel_err(oj.get(oname), gj.get(gname), floor=1e-6)
            assert e.max() <= 1e-4, (oname, gname, float(e.max()))
        log.append("finalize: outputs within 1e-4")
    finally:
        print("\n  " + "\n  ".join(log))
        gj.release()


def test_c1_lockstep_128(engine_ctx):
    """cornell_box at 128x128 with the full-size settings: 16 base spp + 3
    adaptive passes of 4 spp, StatMC radius 10, 3 variance iterations,
    aperture 0.14 / focus 1.05, seed 1234."""
    _lockstep(engine_ctx, ob.config_c1(128, 128), SEED)


def test_c0_lockstep_256(engine_ctx):
    """cornell_box at its full 256x256: 4 + 3x4 spp, default StatMC settings."""
    _lockstep(engine_ctx, ob.config_c0(256, 256), SEED)


def _procedural(level, spheres, seed, w, h):
    pc = eng.procedural_config(seed=seed, subdivision_level=level, n_spheres=spheres)
    rc = ob.RenderConfig(width=w, height=h, spp=8, fov=40.0, cam_pos=(0.0, 0.6, 3.4),
                         cam_dir=(0.0, -0.17, -1.0), focus_distance=3.4, aperture=0.06,
                         adaptive_spp=8, adaptive_passes=4)
    rc.max_depth = 8
    return pc, rc


def test_c2_settings_lockstep_procedural(engine_ctx):
    """The procedural scene's settings (8 base spp + 4 adaptive passes of 8 spp,
    max depth 8, thin lens, environment, GGX spheres, BVH traversal) at a size
    the oracle finishes in seconds."""
    pc, rc = _procedural(4, 48, 7, 96, 64)
    name = "procedural-lockstep-7-4-48"
    engine_ctx.procedural_scene(name, pc)
    _lockstep(engine_ctx, rc, SEED, scene=name, oracle_scene=eng.procedural_arrays(pc), budget="procedural")


@pytest.mark.parametrize("which,w,h", [("c1", 128, 128), ("c0", 256, 256)])
def test_whole_job_matches_oracle(engine_ctx, which, w, h):
    rc = ob.config_c1(w, h) if which == "c1" else ob.config_c0(w, h)
    oj = ob.OracleJob(rc, SEED, nthreads=8, skip_dead=True)
    oj.render()
    gj = eng.Job(engine_ctx, engine_config(rc), engine_camera(rc), SEED, submit=True)
    gj.wait()
    def other():
        o = ob.OracleJob(rc, SEED + 1, nthreads=8, skip_dead=True)
        o.render()
        return o

    try:
        compare_whole_job(oj, gj, f"{which} {w}x{h}", other)
    finally:
        gj.release()


def test_served_stream_matches_oracle():
    """Requests of different sizes, spp, pass counts and scenes (cornell_box
    with and without a lens, the procedural scene with max depth 8) submitted
    at once to the request loop; each finished job is compared with the
    oracle's render() of the same request."""
    ctx = eng.Context(0)
    pc, prc = _procedural(3, 24, 11, 64, 48)
    prc.adaptive_passes = 2
    pname = "procedural-served-11-3-24"
    ctx.procedural_scene(pname, pc)
    reqs = []
    for k, (w, h, spp, passes, ap, radius) in enumerate(
            [(96, 64, 8, 3, 0.14, 10), (64, 48, 4, 0, 0.0, 7), (80, 80, 4, 2, 0.0, 5), (72, 40, 16, 1, 0.14, 10)]):
        rc = ob.config_c0(w, h)
        rc.spp, rc.adaptive_passes, rc.aperture, rc.color_window_radius = spp, passes, ap, radius
        rc.focus_distance = 1.05 if ap > 0 else 1.0
        reqs.append(("cornell_box", rc, None, 2000 + k))
    reqs.insert(2, (pname, prc, eng.procedural_arrays(pc), 3000))
    srv = eng.Server(ctx, max_in_flight=4, graph_passes=True)
    try:
        ids = [srv.submit(engine_config(rc), engine_camera(rc), seed, scene=scene) for scene, rc, _, seed in reqs]
        for (scene, rc, arrays, seed), rid in zip(reqs, ids):
            j, _ = srv.wait(rid)
            oj = ob.OracleJob(rc, seed, nthreads=8, skip_dead=True, scene=arrays)
            oj.render()

            def other(rc=rc, seed=seed, arrays=arrays):
                o = ob.OracleJob(rc, seed + 1, nthreads=8, skip_dead=True, scene=arrays)
                o.render()
                return o

            compare_whole_job(oj, j, f"served {scene} {rc.width}x{rc.height} spp {rc.spp} passes {rc.adaptive_passes}",
                        other)
            j.release()
        st = srv.stats()
        assert st["completed"] == len(reqs) and st["failed"] == 0
        assert st["max_in_flight_seen"] >= 2
    finally:
        srv.close()
        ctx.close()
