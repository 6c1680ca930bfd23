"""Serving: hipGraph-replayed adaptive passes and the request loop.

A job's result must not depend on how it was launched: kernel by kernel or
from a captured graph, alone or admitted together with other jobs whose work
overlaps on the device. Both are compared bit for bit.
"""
import numpy as np
import pytest

import oracle_bind as ob
import rpd_engine as eng
from parity_util import engine_camera, engine_config

pytestmark = pytest.mark.gpu

OUTPUTS = ("radiance", "raw_radiance", "denoised", "variance_denoised", "uncertainty", "sample_count")


def _outputs(job):
    return {n: job.get(n) for n in OUTPUTS}


def _same(a, b, what):
    for n in OUTPUTS:
        np.testing.assert_array_equal(a[n], b[n], err_msg=f"{what}: {n}")


def _requests():
    reqs = []
    for k, (w, h, spp, passes, ap, statmc) in enumerate(
            [(64, 48, 4, 3, 0.0, 1), (96, 64, 8, 2, 0.14, 1), (40, 40, 2, 0, 0.0, 1), (80, 56, 4, 4, 0.14, 1),
             (48, 32, 4, 0, 0.0, 0)]):
        rc = ob.config_c0(w, h)
        rc.spp, rc.adaptive_passes, rc.aperture, rc.focus_distance, rc.use_statmc = spp, passes, ap, 1.05, statmc
        reqs.append((engine_config(rc), engine_camera(rc), 1000 + k))
    return reqs


def test_graph_passes_identical_to_eager():
    ctx = eng.Context(0)
    try:
        res = []
        for graph in (False, True):
            ctx.set_graph_passes(graph)
            outs = []
            for cfg, cam, seed in _requests():
                j = eng.Job(ctx, cfg, cam, seed, submit=True)
                j.wait()
                outs.append(_outputs(j))
                j.release()
            res.append(outs)
        for k, (a, b) in enumerate(zip(*res)):
            _same(a, b, f"request {k} graph vs eager")
    finally:
        ctx.close()


def test_server_results_same_alone_and_in_the_mix():
    ctx = eng.Context(0)
    srv = eng.Server(ctx, max_in_flight=4, graph_passes=True)
    try:
        alone = []
        for cfg, cam, seed in _requests():
            rid = srv.submit(cfg, cam, seed)
            j, _ = srv.wait(rid)
            alone.append(_outputs(j))
            j.release()
        for rep in range(2):  # everything at once: admitted together, overlapping on the device
            ids = [srv.submit(cfg, cam, seed) for cfg, cam, seed in _requests()]
            for k, rid in enumerate(ids):
                j, ms = srv.wait(rid)
                assert ms > 0
                _same(alone[k], _outputs(j), f"request {k} in the mix (round {rep})")
                j.release()
        st = srv.stats()
        assert st["completed"] == 15 and st["failed"] == 0
        assert st["max_in_flight_seen"] >= 2
        assert st["latency_p99_ms"] >= st["latency_p50_ms"] > 0
    finally:
        srv.close()
        ctx.close()


def test_server_admission_by_memory_budget():
    """With room for about one job at a time the loop still finishes all. Held
    bytes count finished jobs until released, so a request the caller waits for
    while a later, finished one is unclaimed goes over the budget by at most
    that one job."""
    ctx = eng.Context(0)
    reqs = _requests()[:3]
    biggest = max(eng.job_bytes(cfg) for cfg, _, _ in reqs)
    srv = eng.Server(ctx, max_in_flight=4, memory_budget=int(biggest * 1.2), graph_passes=False)
    try:
        ids = [srv.submit(cfg, cam, seed) for cfg, cam, seed in reqs]
        for rid in ids:
            j, _ = srv.wait(rid)
            j.release()
        st = srv.stats()
        assert st["completed"] == 3
        assert st["max_in_flight_seen"] <= 2
        assert st["peak_bytes"] <= int(biggest * 1.2) + biggest
        with pytest.raises(eng.EngineError):   # larger than the whole budget: refused at submit
            big = eng.default_config(image_width=1024, image_height=1024, samples_per_pixel=64, use_statmc=1)
            srv.submit(big, eng.make_camera(), 1)
    finally:
        srv.close()
        ctx.close()


def test_server_admits_a_waited_request_while_claimed_jobs_are_held():
    """Claimed jobs hold their device memory until released; with room for one job,
    a caller that keeps every job it claimed must still get the next request
    (admitted over the budget once nothing runs) and the same results."""
    ctx = eng.Context(0)
    reqs = _requests()[:3]
    biggest = max(eng.job_bytes(cfg) for cfg, _, _ in reqs)
    srv = eng.Server(ctx, max_in_flight=4, memory_budget=int(biggest * 1.2), graph_passes=False)
    kept = []
    try:
        alone = []
        for cfg, cam, seed in reqs:
            j = eng.Job(ctx, cfg, cam, seed, submit=True)
            j.wait()
            alone.append(_outputs(j))
            j.release()
        ids = [srv.submit(cfg, cam, seed) for cfg, cam, seed in reqs]
        for k, rid in enumerate(ids):
            j, _ = srv.wait(rid)
            kept.append(j)
            _same(alone[k], _outputs(j), f"request {k} with earlier jobs held")
        st = srv.stats()
        assert st["completed"] == 3 and st["failed"] == 0
        assert st["peak_bytes"] >= sum(eng.job_bytes(cfg) for cfg, _, _ in reqs[:2]) > int(biggest * 1.2)
    finally:
        for j in kept:
            j.release()
        srv.close()
        ctx.close()


def test_reused_job_buffers_give_the_same_results():
    """Released jobs' device buffers are kept by the context and handed to later
    jobs of the same sizes. Each request runs after a job of the same sizes but
    another seed, so it reuses buffers holding other data, and must match the
    run on fresh buffers bit for bit."""
    ctx = eng.Context(0)
    try:
        assert ctx.set_buffer_cache(0) == 0
        fresh = []
        for cfg, cam, seed in _requests():
            j = eng.Job(ctx, cfg, cam, seed, submit=True)
            j.wait()
            fresh.append(_outputs(j))
            j.release()
        assert ctx.set_buffer_cache(1 << 34) == 0
        for k, (cfg, cam, seed) in enumerate(_requests()):
            other = eng.Job(ctx, cfg, cam, seed + 777, submit=True)
            other.wait()
            held = ctx.memory()["in_use"]
            other.release()
            kept = ctx.set_buffer_cache(1 << 34)
            assert kept >= held // 2
            j = eng.Job(ctx, cfg, cam, seed, submit=True)
            j.wait()
            _same(fresh[k], _outputs(j), f"request {k} on reused buffers")
            j.release()
        assert ctx.set_buffer_cache(0) == 0
    finally:
        ctx.close()
