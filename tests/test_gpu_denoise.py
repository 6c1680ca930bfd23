"""Denoising given per-sample records on the GPU vs. the CPU oracle, and the
batched denoise (many uneven images per launch) vs. the same images alone.

The oracle folds the same records in the reference's order (pixel-major,
samples in order, one thread), so pixel statistics, features and the RPF grid
are compared bit for bit; the StatMC result has no pixel beyond 1e-4 relative
(measured: bit-identical), alpha within 1e-6 and uncertainty within 1e-5.
"""
import numpy as np
import pytest

import oracle_bind as ob
import rpd_engine as eng
from parity_util import engine_config, rel_err

pytestmark = pytest.mark.gpu

SEED = 1234


def _denoiser(ctx, rc, spp):
    return eng.Job(ctx, engine_config(rc), None, 0, denoiser_spp=spp)


def _compare_state(oj, gj, what):
    np.testing.assert_array_equal(oj.get("stats").view(np.int32), gj.get("stats").view(np.int32), err_msg=what)
    np.testing.assert_array_equal(oj.get("hit_count"), gj.get("hit_count"), err_msg=what)
    np.testing.assert_array_equal(oj.get("sample_count"), gj.get("sample_count"), err_msg=what)
    np.testing.assert_array_equal(oj.get("grid"), gj.get("grid"), err_msg=what)


def _compare_denoised(oj, gj, what):
    e = rel_err(oj.get("denoised"), gj.get("denoised"), floor=1e-6)
    bad = (e > 1e-4).any(axis=1)
    print(what, "denoised pixels beyond 1e-4:", int(bad.sum()), "of", bad.size, "max", float(e.max()))
    assert not bad.any()
    np.testing.assert_allclose(oj.get("alpha"), gj.get("alpha"), rtol=0, atol=1e-6)
    assert rel_err(oj.get("uncertainty"), gj.get("uncertainty"), floor=1e-12).max() <= 1e-5


@pytest.mark.parametrize("aperture", [0.0, 0.14])
def test_traced_records_denoise_like_oracle(engine_ctx, aperture):
    rc = ob.config_c0(48, 40)
    rc.aperture = aperture
    rc.focus_distance = 1.05
    oj = ob.OracleJob(rc, SEED, nthreads=1, keep_records=True)
    counts = np.full(rc.width * rc.height, rc.spp, np.int32)
    oj.run_pass(counts, statmc=True)
    rows = oj.get("records")
    gj = _denoiser(engine_ctx, rc, rc.spp)
    gj.load_records(counts, eng.records_from_rows(rows))
    gj.denoise_records()
    _compare_state(oj, gj, "traced records")
    oj.recompute_sensitivity()
    oj.statmc()
    _compare_denoised(oj, gj, "traced records")


# grid steps 4, 8, 16: each gather support-window size; steps 20 and 35 walk
# their windows (up to 40x40 and 70x70 positions) in segments of 1024
@pytest.mark.parametrize("tile,size", [(8, (37, 29)), (16, (37, 29)), (32, (37, 29)), (40, (53, 47)),
                                       (70, (80, 75))])
def test_uneven_counts_and_nonfinite_records(engine_ctx, tile, size):
    rc = ob.config_c0(*size)
    rc.rpf_tile_size = tile
    rng = np.random.default_rng(5)
    counts = rng.integers(0, 9, rc.width * rc.height).astype(np.int32)
    n = int(counts.sum())
    rows = np.zeros((n, 32), np.float32)
    rows[:, 0] = np.repeat(np.arange(counts.size), counts)
    rows[:, 1:4] = rng.lognormal(-1.0, 1.5, (n, 3))
    rows[rng.random(n) < 0.01, 1] = np.inf                      # dropped like non-finite traced samples
    rows[:, 4] = np.isfinite(rows[:, 1:4]).all(axis=1)
    hit = rng.random(n) < 0.9
    rows[:, 5] = hit
    nrm = rng.normal(size=(n, 3))
    nrm[:, 2] = np.abs(nrm[:, 2]) + 1.0
    rows[:, 6:9] = nrm / np.linalg.norm(nrm, axis=1, keepdims=True)
    rows[:, 9:12] = rng.uniform(0.1, 0.9, (n, 3))
    rows[:, 15] = rng.uniform(1, 20, n)
    rows[~hit, 6:16] = 0.0
    rows[:, 16:26] = rng.random((n, 10))
    rows[:, 26] = rng.integers(0, 20, n)                          # indices above 15 clamp to the last bin
    rows[:, 27] = rng.random(n) < 0.8
    rows[:, 28] = rng.integers(0, 128, n)
    oj = ob.OracleJob(rc, SEED)
    oj.accumulate_records(counts, rows)
    gj = _denoiser(engine_ctx, rc, 8)
    gj.load_records(counts, eng.records_from_rows(rows))
    gj.denoise_records()
    _compare_state(oj, gj, "random records")
    oj.recompute_sensitivity()
    oj.statmc()
    _compare_denoised(oj, gj, "random records")


def test_synthetic_records_denoise_like_oracle(engine_ctx):
    rc = ob.config_c0(64, 48)
    gj = _denoiser(engine_ctx, rc, 12)
    gj.synthesize_records(eng.synth_config(seed=3, samples_per_pixel=12, region_size=8))
    rows = gj.get("records")
    assert rows.shape[0] == 64 * 48 * 12
    counts = np.full(64 * 48, 12, np.int32)
    oj = ob.OracleJob(rc, SEED)
    oj.accumulate_records(counts, rows)
    gj.denoise_records()
    _compare_state(oj, gj, "synthetic records")
    oj.recompute_sensitivity()
    oj.statmc()
    _compare_denoised(oj, gj, "synthetic records")


def test_synthetic_record_statistics(engine_ctx):
    """The generator's stated distribution: lognormal(-1, 1.5) radiance with 2 %
    fireflies x100, dims uniform, three dims at copula correlation 0.6."""
    rc = ob.config_c0(256, 128)
    gj = _denoiser(engine_ctx, rc, 16)
    gj.synthesize_records(eng.synth_config(seed=11, samples_per_pixel=16, region_size=16))
    rows = gj.get("records")
    from scipy.stats import norm
    # log radiance is a mixture: N(mu, sigma) and, with p = 0.02, the same shifted by ln(100)
    lum = rows[:, 1:4].astype(np.float64).mean(axis=1)             # the tint keeps the channel mean
    z = (np.log(lum) + 1.0) / 1.5
    p, shift = 0.02, np.log(100.0) / 1.5
    sd = np.sqrt(1.0 + p * (1 - p) * shift ** 2)
    assert abs(z.mean() - p * shift) < 0.01, z.mean()
    assert abs(z.std() - sd) < 0.01, z.std()
    tail = (1 - p) * norm.sf(3.0) + p * norm.sf(3.0 - shift)
    assert abs((z > 3.0).mean() - tail) < 0.001, ((z > 3.0).mean(), tail)
    dims = rows[:, 18:26]
    assert (dims >= 0).all() and (dims < 1).all()
    assert np.all(np.abs(dims.mean(axis=0) - 0.5) < 0.01)
    for col in (18, 20, 22):                                         # lens x, brdf x, light x: copula rho 0.6
        g = norm.ppf(np.clip(rows[:, col], 1e-7, 1 - 1e-7))
        r = np.corrcoef(z, g)[0, 1]
        assert abs(r - 0.6 / sd) < 0.01, (col, r, 0.6 / sd)
    for col in (19, 21, 23, 24):
        g = norm.ppf(np.clip(rows[:, col], 1e-7, 1 - 1e-7))
        assert abs(np.corrcoef(z, g)[0, 1]) < 0.01
    nrm = rows[:, 6:9]
    assert np.allclose(np.linalg.norm(nrm, axis=1), 1.0, atol=1e-5)
    assert (rows[:, 15] >= 1).all() and (rows[:, 15] <= 20).all()


def test_batch_identical_to_alone(engine_ctx):
    """Uneven images, sample counts and grid steps in one batch give the same bits as each
    image alone (alone, each runs the gather built for its own step; in the batch, all run the
    one built for the largest; a step above 16 walks each window in segments)."""
    shapes = [(64, 48, 4), (100, 37, 7), (33, 70, 16), (128, 128, 2), (17, 9, 5), (90, 60, 3)]
    tiles = [8, 16, 6, 32, 8, 44]
    outs = []
    for batched in (False, True):
        jobs = []
        for k, (w, h, spp) in enumerate(shapes):
            rc = ob.config_c0(w, h)
            rc.color_window_radius = 3 + k
            rc.rpf_tile_size = tiles[k]
            j = _denoiser(engine_ctx, rc, spp)
            j.synthesize_records(eng.synth_config(seed=100 + k, samples_per_pixel=spp, region_size=4 + 3 * k))
            jobs.append(j)
        if batched:
            ms = eng.denoise_batch(engine_ctx, jobs)
            assert ms > 0
        else:
            for j in jobs:
                j.denoise_records()
        outs.append([{n: j.get(n) for n in ("denoised", "uncertainty", "alpha", "grid", "tiles", "stats")}
                     for j in jobs])
        for j in jobs:
            j.release()
    for a, b in zip(*outs):
        for n in a:
            np.testing.assert_array_equal(a[n], b[n], err_msg=n)
