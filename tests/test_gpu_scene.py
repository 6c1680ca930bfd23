"""The procedural scene on the GPU (BVH traversal, GGX spheres, four area
lights, environment, thin lens, max depth 8) against the oracle tracing the
same scene arrays: per-sample records of identical random streams.
"""
import numpy as np
import pytest

import oracle_bind as ob
import rpd_engine as eng
from parity_util import engine_camera, engine_config, rel_err

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
    assert orec.shape == grec.shape
    np.testing.assert_array_equal(orec[:, 0], grec[:, 0])
    np.testing.assert_array_equal(orec[:, 16:20], grec[:, 16:20])
    rad_bad = (rel_err(orec[:, 1:4], grec[:, 1:4]) > 1e-4).any(axis=1)
    dims_bad = (orec[:, 20:29] != grec[:, 20:29]).any(axis=1)
    flag_bad = (orec[:, 4:6] != grec[:, 4:6]).any(axis=1)
    feat_bad = (rel_err(orec[:, 6:16], grec[:, 6:16], floor=1.0) > 1e-5).any(axis=1)
    bad = rad_bad | dims_bad | flag_bad | feat_bad
    print(name, "samples", len(bad), "outliers", int(bad.sum()), "(radiance", int(rad_bad.sum()), "dims",
          int(dims_bad.sum()), "flags", int(flag_bad.sum()), "features", int(feat_bad.sum()), ")",
          "env captured", int((orec[:, 28].astype(int) & 32 != 0).sum()), "hits", int(orec[:, 5].sum()))
    assert bad.mean() <= 5e-3
    assert (orec[:, 28].astype(int) & 32 != 0).any()      # the environment dimension is exercised
    assert rel_err(orec[~bad, 1:4], grec[~bad, 1:4]).max() <= 1e-4
