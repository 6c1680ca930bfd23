"""StatMC window kernel tile shapes and their tuner.

The window kernel (main.cpp:998-1127 per pixel) runs over tiles of one of
several shapes; the context times the shapes on the first launch of each
(radius, image-size class) and keeps the fastest. Every shape must give the
same bits: per-pixel arithmetic does not depend on the tile, and the staged
(LDS) and unstaged gate inputs are the same values. Each shape is compared bit
for bit with the default 16x16 shape, for radii whose tile-plus-halo fits the
staging LDS and for radius 24, whose halo does not; the default is compared
with the oracle (no pixel beyond 1e-4).
"""
import numpy as np
import pytest

import oracle_bind as ob
import rpd_engine as eng
from parity_util import engine_camera, engine_config, rel_err

pytestmark = pytest.mark.gpu

OUTS = ("denoised", "alpha", "uncertainty", "var_total", "var_eff", "var_ratio")


def _inject(oj, gj):
    gj.set("stats", oj.get("stats"))
    sums = np.concatenate([oj.get("radiance"), oj.get("albedo"), oj.get("normal"), oj.get("world_pos"),
                           oj.get("depth")[:, None]], axis=1)
    gj.set("sums", sums)
    gj.set("hit_count", oj.get("hit_count"))
    gj.set("sample_count", oj.get("sample_count"))
    gj.set("tiles", oj.get("tiles").ravel())
    gj.set("sensitivity_gradient", oj.get("sensitivity_gradient"))


@pytest.mark.parametrize("radius", [3, 7, 10, 24])
def test_every_tile_shape_gives_the_same_bits(radius):
    rc = ob.config_c0(75, 53)
    rc.color_window_radius = radius
    oj = ob.OracleJob(rc, 1234, nthreads=4)
    oj.run_pass(np.full(rc.width * rc.height, rc.spp, np.int32), statmc=True)
    oj.recompute_sensitivity()
    oj.statmc()
    ctx = eng.Context(0)
    try:
        gj = eng.Job(ctx, engine_config(rc), engine_camera(rc), 1234)
        _inject(oj, gj)
        shapes = eng.statmc_tile_shapes()
        assert shapes[0] == (16, 16) and len(shapes) >= 4
        res = {}
        for w, h in shapes:
            ctx.set_statmc_tile(w, h)
            gj.statmc()
            res[(w, h)] = {k: gj.get(k) for k in OUTS}
        base = res[(16, 16)]
        e = rel_err(oj.get("denoised"), base["denoised"]).max(axis=1)
        print(f"radius {radius}: default shape vs oracle max rel err {e.max():.2e}")
        assert not (e > 1e-4).any()
        for shape, r in res.items():
            for k in OUTS:
                np.testing.assert_array_equal(r[k], base[k], err_msg=f"{shape} {k}")
        # tuned: times every shape, keeps one, same bits
        ctx.set_statmc_tile(0, 0)
        assert ctx.statmc_tile(radius, rc.width * rc.height) == (0, 0)
        gj.statmc()
        chosen = ctx.statmc_tile(radius, rc.width * rc.height)
        print(f"radius {radius}: tuned shape {chosen}")
        assert chosen in shapes
        for k in OUTS:
            np.testing.assert_array_equal(gj.get(k), base[k], err_msg=f"tuned {k}")
        gj.statmc()                                   # the tuned shape, reused
        np.testing.assert_array_equal(gj.get("denoised"), base["denoised"])
        gj.release()
    finally:
        ctx.close()


def test_batched_denoise_same_bits_for_every_tile_shape():
    ctx = eng.Context(0)
    try:
        def run(shape):
            if shape is not None:
                ctx.set_statmc_tile(*shape)
            jobs = []
            for k, (w, h, spp, r) in enumerate([(40, 30, 4, 3), (97, 61, 8, 10), (33, 70, 6, 7)]):
                cfg = eng.default_config(image_width=w, image_height=h, samples_per_pixel=spp, use_statmc=1,
                                         color_window_radius=r)
                j = eng.Job(ctx, cfg, None, 0, denoiser_spp=spp)
                j.synthesize_records(eng.synth_config(seed=50 + k, samples_per_pixel=spp))
                jobs.append(j)
            eng.denoise_batch(ctx, jobs)
            out = [j.get("denoised") for j in jobs]
            for j in jobs:
                j.release()
            return out

        base = run((16, 16))
        for shape in eng.statmc_tile_shapes()[1:] + [(0, 0)]:
            for a, b in zip(base, run(shape)):
                np.testing.assert_array_equal(a, b, err_msg=str(shape))
    finally:
        ctx.close()
