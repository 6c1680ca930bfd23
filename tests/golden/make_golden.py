#!/usr/bin/env python3
"""Writes the small golden fixtures of tests/golden/ from the CPU oracle.

Each fixture holds inputs and the oracle's outputs for one step of the
render -> RPF grid -> StatMC -> allocation path at a size the oracle
finishes in well under a second. The CPU tests check that the oracle still
reproduces them bit for bit; the GPU tests compare the engine with them.

Run from the repository root after building the oracle:
    python3 tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
import oracle_bind as ob  # noqa: E402


def cornell_pass():
    rc = ob.config_c0(16, 12)
    rc.spp = 2
    rc.aperture = 0.14
    rc.focus_distance = 1.05
    oj = ob.OracleJob(rc, 1234, nthreads=1, keep_records=True)
    counts = np.full(16 * 12, rc.spp, np.int32)
    oj.run_pass(counts, statmc=True)
    rec = oj.get("records")
    oj.recompute_sensitivity()
    oj.statmc()
    return dict(width=16, height=12, spp=2, aperture=0.14, focus_distance=1.05, seed=1234,
                records=rec, stats=oj.get("stats"), grid=oj.get("grid"), denoised=oj.get("denoised"),
                alpha=oj.get("alpha"))


def allocation():
    rc = ob.config_c0(24, 16)
    rng = np.random.default_rng(3)
    imp = rng.lognormal(-1.0, 1.2, 24 * 16).astype(np.float32)
    imp[rng.random(imp.size) < 0.1] = 0.0
    oj = ob.OracleJob(rc, 99)
    return dict(width=24, height=16, seed=99, importance=imp, counts=oj.counts_from_importance(imp))


def full_job():
    rc = ob.config_c0(12, 10)
    rc.spp = 2
    rc.adaptive_spp = 2
    rc.adaptive_passes = 2
    oj = ob.OracleJob(rc, 5, nthreads=1, skip_dead=True)
    oj.render()
    return dict(width=12, height=10, spp=2, adaptive_spp=2, adaptive_passes=2, seed=5,
                radiance=oj.get("radiance"), raw_radiance=oj.get("raw_radiance"),
                sample_count=oj.get("sample_count"), variance_denoised=oj.get("variance_denoised"))


def main():
    for name, fn in (("cornell_pass_16x12", cornell_pass), ("allocation_24x16", allocation),
                     ("cornell_job_12x10", full_job)):
        d = fn()
        np.savez_compressed(os.path.join(HERE, name + ".npz"), **d)
        print(name, {k: getattr(v, "shape", v) for k, v in d.items()})


if __name__ == "__main__":
    main()
