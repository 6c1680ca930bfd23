"""GPU engine vs. the CPU oracle on the same seeded inputs.

Each stage test feeds the engine the oracle's own state (so both sides start
from identical bits) and compares the stage output; the trace tests compare
per-sample records of identical RNG streams. Tolerances: random draws, bin
inputs, flags and sample counts are compared exactly; radiance within 1e-4
relative per sample, with samples whose path took a different branch at an
intersection rounding boundary counted as outliers and held to
RECORD_OUTLIER_BUDGET; StatMC and variance smoothing from identical inputs have
no pixel beyond 1e-4; allocation counts are exact.
"""
import numpy as np
import pytest

import oracle_bind as ob
import rpd_engine as eng
from parity_util import RECORD_OUTLIER_BUDGET, check_records, compare_whole_job, engine_camera, engine_config, rel_err

BUDGET = RECORD_OUTLIER_BUDGET["cornell_box"]

pytestmark = pytest.mark.gpu

SEED = 1234


def _pair(rc, seed=SEED):
    oj = ob.OracleJob(rc, seed, nthreads=1, keep_records=True)
    return oj


def _gpu_job(ctx, rc, seed=SEED):
    return eng.Job(ctx, engine_config(rc), engine_camera(rc), seed)


@pytest.mark.parametrize("aperture", [0.0, 0.14])
def test_trace_records_match_oracle(engine_ctx, aperture):
    rc = ob.config_c0(40, 32)
    rc.aperture = aperture
    rc.focus_distance = 1.05 if aperture > 0 else 1.0
    oj = _pair(rc)
    oj.run_pass(np.full(40 * 32, rc.spp, np.int32), statmc=True)
    gj = _gpu_job(engine_ctx, rc)
    gj.run_pass(None, statmc=True)
    check_records(oj.get("records"), gj.get("records"), BUDGET, f"aperture {aperture}:")


def test_trace_records_past_the_first_twist_match_oracle(engine_ctx):
    """32 samples per pixel draw about a thousand numbers per pixel: tape words
    227-623 (whose twist reads back words the tape kernel wrote earlier) and the
    in-kernel re-twist after word 623 are all consumed."""
    rc = ob.config_c0(16, 12)
    rc.spp = 32
    oj = _pair(rc)
    oj.run_pass(np.full(16 * 12, rc.spp, np.int32), statmc=True)
    gj = _gpu_job(engine_ctx, rc)
    gj.run_pass(None, statmc=True)
    orec, grec = oj.get("records"), gj.get("records")
    assert len(orec) == 16 * 12 * 32
    check_records(orec, grec, BUDGET, "32 spp:")


def test_trace_records_past_the_next_state_prefix_match_oracle(engine_ctx):
    """64 samples per pixel draw about 1,400 numbers per pixel: past word 623
    into the next state's first 128 words that the tape kernel wrote, past
    those (the tape is then twisted in place and the draws go on from word
    128), and past the end of that state (twisted again, from word 0)."""
    rc = ob.config_c0(12, 10)
    rc.spp = 64
    oj = _pair(rc)
    oj.run_pass(np.full(12 * 10, rc.spp, np.int32), statmc=True)
    gj = _gpu_job(engine_ctx, rc)
    gj.run_pass(None, statmc=True)
    orec, grec = oj.get("records"), gj.get("records")
    assert len(orec) == 12 * 10 * 64
    check_records(orec, grec, BUDGET, "64 spp:")


@pytest.mark.parametrize("per_sample,margin", [(48, 64), (4, 8), (0, 0), (0, 624)])
def test_trace_records_with_short_tape_prefixes_match_oracle(engine_ctx, per_sample, margin):
    """Tapes written only in part before the pass (48 words per sample and 64
    more, the default; 4 and 8, so that nearly every pixel's lane extends its
    own tape; none at all; and the whole 624 words) give the same draws, on
    uneven counts of 1 to 12 samples per pixel and a second pass seeded past
    the first one's samples."""
    rc = ob.config_c0(24, 20)
    rc.spp = 12                       # record room for up to 12 samples per pixel
    rng = np.random.default_rng(per_sample + margin)
    counts = rng.integers(1, 13, rc.width * rc.height).astype(np.int32)
    oj = _pair(rc)
    gj = _gpu_job(engine_ctx, rc)
    engine_ctx.set_tape_prefix(per_sample, margin)
    try:
        oj.run_pass(np.full(rc.width * rc.height, 2, np.int32), statmc=True)
        gj.run_pass(np.full(rc.width * rc.height, 2, np.int32), statmc=True)
        oj.run_pass(counts, statmc=True)
        gj.run_pass(counts, statmc=True)
        check_records(oj.get("records"), gj.get("records"), BUDGET, f"prefix {per_sample}/{margin}:")
    finally:
        engine_ctx.set_tape_prefix(48, 64)
        gj.release()


def test_trace_uniform_render_matches_oracle(engine_ctx):
    rc = ob.config_c0(24, 20)
    rc.use_statmc = 0
    oj = _pair(rc)
    oj.run_pass(np.full(24 * 20, rc.spp, np.int32), statmc=False)
    gj = _gpu_job(engine_ctx, rc)
    gj.run_pass(None, statmc=False)
    check_records(oj.get("records"), gj.get("records"), BUDGET, "uniform render:")


def test_pass_state_matches_oracle(engine_ctx):
    rc = ob.config_c0(40, 32)
    oj = _pair(rc)
    oj.run_pass(np.full(40 * 32, rc.spp, np.int32), statmc=True)
    gj = _gpu_job(engine_ctx, rc)
    gj.run_pass(None, statmc=True)
    bad = check_records(oj.get("records"), gj.get("records"), BUDGET, "pass state:")
    out_px = np.zeros(rc.width * rc.height, bool)
    out_px[oj.get("records")[bad, 0].astype(np.int64)] = True
    np.testing.assert_array_equal(oj.get("hit_count"), gj.get("hit_count"))
    np.testing.assert_array_equal(oj.get("sample_count"), gj.get("sample_count"))
    os_, gs = oj.get("stats"), gj.get("stats")
    np.testing.assert_array_equal(os_[:, 2].view(np.int32), gs[:, 2].view(np.int32))
    cols = [0, 3, 4, 5, 6, 7, 8]
    px_bad = (rel_err(os_[:, cols], gs[:, cols], floor=1e-3) > 1e-4).any(axis=1)
    print("pixels with stat differences:", px_bad.sum(), "outlier pixels:", out_px.sum())
    assert not (px_bad & ~out_px).any()
    # grid accumulators are sums of non-negative per-sample terms: within the
    # per-sample tolerance, except nodes an outlier splatted into (at most 4 each)
    og, gg = oj.get("grid"), gj.get("grid")
    assert og.shape == gg.shape
    node_err = (np.abs(og - gg) / np.maximum(np.abs(og), 1e-300)).max(axis=1)
    node_err[np.abs(og - gg).max(axis=1) == 0.0] = 0.0
    print("grid max rel err", node_err.max(), "nodes beyond 1e-4:", (node_err > 1e-4).sum(), "of", len(og))
    assert (node_err > 1e-4).sum() <= 4 * bad.sum()


def _oracle_after_base(rc, seed=SEED):
    oj = ob.OracleJob(rc, seed, nthreads=1)
    oj.run_pass(np.full(rc.width * rc.height, rc.spp, np.int32), statmc=True)
    return oj


def _inject_pixel_state(oj, gj):
    gj.set("stats", oj.get("stats"))
    sums = np.concatenate([oj.get("radiance"), oj.get("albedo"), oj.get("normal"), oj.get("world_pos"),
                           oj.get("depth")[:, None]], axis=1)
    gj.set("sums", sums)
    gj.set("hit_count", oj.get("hit_count"))
    gj.set("sample_count", oj.get("sample_count"))


def test_sensitivity_stage_identical_inputs(engine_ctx):
    rc = ob.config_c0(48, 40)
    oj = _oracle_after_base(rc)
    oj.recompute_sensitivity()
    gj = _gpu_job(engine_ctx, rc)
    _inject_pixel_state(oj, gj)
    gj.set("grid", oj.get("grid"))
    gj.recompute_sensitivity()
    ot, gt = oj.get("tiles"), gj.get("tiles")
    err = np.abs(ot - gt)
    print("tile max abs err", err.max())
    assert err.max() <= 1e-5
    np.testing.assert_allclose(gj.get("sensitivity_gradient"), oj.get("sensitivity_gradient"), atol=1e-5)
    np.testing.assert_allclose(gj.get("sensitivity_confidence"), oj.get("sensitivity_confidence"), atol=1e-5)


@pytest.mark.parametrize("radius,alpha", [(7, 0.05), (10, 0.01), (3, 0.2)])
def test_statmc_stage_identical_inputs(engine_ctx, radius, alpha):
    rc = ob.config_c0(48, 40)
    rc.color_window_radius = radius
    rc.color_compat_alpha = alpha
    oj = _oracle_after_base(rc)
    oj.recompute_sensitivity()
    oj.statmc()
    gj = _gpu_job(engine_ctx, rc)
    _inject_pixel_state(oj, gj)
    gj.set("tiles", oj.get("tiles"))
    gj.set("sensitivity_gradient", oj.get("sensitivity_gradient"))
    gj.statmc()
    od, gd = oj.get("denoised"), gj.get("denoised")
    e = rel_err(od, gd)
    print("denoised max rel err", e.max(), "pixels > 1e-4:", (e > 1e-4).any(axis=1).sum(),
          "bit-identical pixels:", (od == gd).all(axis=1).mean())
    assert not (e > 1e-4).any()
    np.testing.assert_allclose(gj.get("alpha"), oj.get("alpha"), rtol=0, atol=1e-6)
    assert rel_err(oj.get("uncertainty"), gj.get("uncertainty"), floor=1e-12).max() <= 1e-5
    for name in ("var_total", "var_eff", "var_ratio"):
        np.testing.assert_array_equal(oj.get(name), gj.get(name))


@pytest.mark.parametrize("radius", [3, 10])
def test_variance_stage_identical_inputs(engine_ctx, radius):
    """Bit-exact from identical statistics; radius 3 filters from the tile
    staged in LDS, radius 10 (past its halo limit) from global memory."""
    rc = ob.config_c0(48, 40)
    rc.var_window_radius = radius
    rc.var_iterations = 6
    rc.var_compat_sigma = 3.0
    rc.var_shrinkage_k = 0.7
    oj = _oracle_after_base(rc)
    oj.variance()
    gj = _gpu_job(engine_ctx, rc)
    _inject_pixel_state(oj, gj)
    gj.variance()
    np.testing.assert_array_equal(oj.get("variance_denoised"), gj.get("variance_denoised"))


@pytest.mark.parametrize("base_samples,radius,spp", [(0, 1, 4), (2, 1, 4), (1, 0, 4), (0, 2, 1), (0, 9, 4)])
def test_allocation_from_state(engine_ctx, base_samples, radius, spp):
    """compute_adaptive_sample_counts from identical statistics: importance
    agrees to one float ulp (the smoothing's expf is correctly rounded here and
    glibc's is not on ~1e-4 of arguments) and the counts agree exactly."""
    rc = ob.config_c0(48, 40)
    rc.adaptive_base_samples = base_samples
    rc.adaptive_importance_smoothing_radius = radius
    rc.adaptive_spp = spp
    oj = _oracle_after_base(rc)
    oj.variance()
    oc = oj.adaptive_counts()
    gj = _gpu_job(engine_ctx, rc)
    _inject_pixel_state(oj, gj)
    gj.set("variance_denoised", oj.get("variance_denoised"))
    gc = gj.adaptive_counts()
    oi, gi = oj.get("adaptive_importance"), gj.get("adaptive_importance")
    ulps = np.abs(oi.view(np.int32).astype(np.int64) - gi.view(np.int32).astype(np.int64))
    print("importance values differing:", int((ulps > 0).sum()), "max ulps", int(ulps.max()))
    assert ulps.max() <= 1
    if radius == 0:
        assert (ulps == 0).all()
    np.testing.assert_array_equal(oc, gc)
    assert gc.sum() <= rc.adaptive_spp * rc.width * rc.height


def _importance_fields(n, rng):
    yield "lognormal", np.exp(rng.normal(-1.0, 1.0, n)).astype(np.float32)
    spiky = np.full(n, 1e-4, np.float32)            # a few pixels far above the cap force clipping sweeps
    spiky[rng.choice(n, 12, replace=False)] = 3.0
    yield "spiky", spiky
    steps = np.repeat(rng.uniform(1e-4, 3.0, n // 16 + 1), 16)[:n].astype(np.float32)
    yield "piecewise", steps
    yield "constant", np.full(n, 0.37, np.float32)
    ties = (np.arange(n) % 7 == 0).astype(np.float32) * 2.5 + np.float32(2.0 ** -20)
    yield "ties", ties
    # heavy tail: each sweep's clips raise the share of the rest, so clips cascade over several sweeps
    yield "pareto", (rng.pareto(0.6, n) + 1e-3).astype(np.float32)
    # a long run of zeros first (pixels without importance), then zeros among the rest
    lead = np.exp(rng.normal(-1.0, 1.0, n)).astype(np.float32)
    lead[: n // 3] = 0.0
    lead[rng.random(n) < 0.2] = 0.0
    yield "leading_zeros", lead


@pytest.mark.parametrize("base_samples,spp,w,h", [(0, 4, 64, 48), (1, 4, 64, 48), (0, 1, 64, 48), (3, 8, 64, 48),
                                                  (0, 4, 640, 480), (1, 4, 1024, 700)])
def test_allocation_bit_exact_from_importance(engine_ctx, base_samples, spp, w, h):
    """main.cpp:1373-1438 on identical importance fields: counts bit-exact."""
    rc = ob.config_c0(w, h)
    rc.adaptive_base_samples = base_samples
    rc.adaptive_spp = spp
    oj = ob.OracleJob(rc, 4321, nthreads=1)
    gj = eng.Job(engine_ctx, engine_config(rc), engine_camera(rc), 4321)
    rng = np.random.default_rng(11 + w)
    for name, imp in _importance_fields(w * h, rng):
        oc = oj.counts_from_importance(imp)
        gc = gj.counts_from_importance(imp)
        assert np.array_equal(oc, gc), (name, int((oc != gc).sum()))
        assert gc.sum() <= spp * w * h


def test_full_pipeline_c0_small(engine_ctx):
    rc = ob.config_c0(48, 40)
    oj = ob.OracleJob(rc, SEED, nthreads=4, skip_dead=True)
    oj.render()
    gj = eng.Job(engine_ctx, engine_config(rc), engine_camera(rc), SEED, submit=True)
    gj.wait()

    def other():
        o = ob.OracleJob(rc, SEED + 1, nthreads=4, skip_dead=True)
        o.render()
        return o

    compare_whole_job(oj, gj, "c0 48x40", other)
