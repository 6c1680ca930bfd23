"""Committed golden fixtures (tests/golden/*.npz, written by
tests/golden/make_golden.py from the oracle).

CPU: the oracle reproduces every fixture bit for bit, so a change to the
oracle that moves any value is caught. GPU: the engine against the same
fixtures, to the parity tolerances (records per sample within 1e-4 relative,
counts exact, StatMC within 1e-4 given the same pass).
"""
import os

import numpy as np
import pytest

import oracle_bind as ob
from parity_util import rel_err

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    with np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def _cornell_rc(g):
    rc = ob.config_c0(int(g["width"]), int(g["height"]))
    rc.spp = int(g["spp"])
    if "aperture" in g:
        rc.aperture = float(g["aperture"])
        rc.focus_distance = float(g["focus_distance"])
    return rc


def test_oracle_reproduces_cornell_pass():
    g = load("cornell_pass_16x12")
    rc = _cornell_rc(g)
    oj = ob.OracleJob(rc, int(g["seed"]), nthreads=1, keep_records=True)
    oj.run_pass(np.full(rc.width * rc.height, rc.spp, np.int32), statmc=True)
    assert np.array_equal(oj.get("records"), g["records"])
    assert np.array_equal(oj.get("stats").view(np.uint32), g["stats"].view(np.uint32))
    assert np.array_equal(oj.get("grid"), g["grid"])
    oj.recompute_sensitivity()
    oj.statmc()
    assert np.array_equal(oj.get("denoised"), g["denoised"])


def test_oracle_reproduces_allocation():
    g = load("allocation_24x16")
    oj = ob.OracleJob(ob.config_c0(int(g["width"]), int(g["height"])), int(g["seed"]))
    assert np.array_equal(oj.counts_from_importance(g["importance"]), g["counts"])
    assert g["counts"].sum() <= 4 * g["importance"].size


def test_oracle_reproduces_full_job():
    g = load("cornell_job_12x10")
    rc = _cornell_rc(g)
    rc.adaptive_spp = int(g["adaptive_spp"])
    rc.adaptive_passes = int(g["adaptive_passes"])
    oj = ob.OracleJob(rc, int(g["seed"]), nthreads=1, skip_dead=True)
    oj.render()
    for k in ("radiance", "raw_radiance", "sample_count", "variance_denoised"):
        assert np.array_equal(oj.get(k), g[k]), k


@pytest.mark.gpu
def test_engine_against_golden_fixtures(engine_ctx):
    import rpd_engine as eng
    from parity_util import RECORD_OUTLIER_BUDGET, check_records, engine_camera, engine_config
    g = load("cornell_pass_16x12")
    rc = _cornell_rc(g)
    gj = eng.Job(engine_ctx, engine_config(rc), engine_camera(rc), int(g["seed"]))
    gj.run_pass(None, statmc=True)
    rec = gj.get("records")
    assert np.array_equal(rec[:, 16:20], g["records"][:, 16:20])            # camera draws
    bad = check_records(g["records"], rec, RECORD_OUTLIER_BUDGET["cornell_box"], "golden pass:")
    gj.recompute_sensitivity()
    gj.statmc()
    e = rel_err(gj.get("denoised"), g["denoised"], floor=1e-6)
    if not bad.any():   # identical records: the denoised image within 1e-4 everywhere
        assert not (e > 1e-4).any()
    a = load("allocation_24x16")
    aj = eng.Job(engine_ctx, engine_config(ob.config_c0(24, 16)), engine_camera(ob.config_c0(24, 16)), int(a["seed"]))
    assert np.array_equal(aj.counts_from_importance(a["importance"]), a["counts"])
