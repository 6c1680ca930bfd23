"""Serving: hipGraph-replayed adaptive passes and the request loop.

A job's result must not depend on how it was launched: kernel by kernel or
from a captured graph, alone or admitted together with other jobs whose work
overlaps on the device, with its denoise tail on its own stream or batched
with other jobs' tails. All are compared bit for bit.
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


@pytest.mark.parametrize("cobatch", [False, True])
def test_server_results_same_alone_and_in_the_mix(cobatch):
    ctx = eng.Context(0)
    plain = []
    for cfg, cam, seed in _requests():
        j = eng.Job(ctx, cfg, cam, seed, submit=True)
        j.wait()
        plain.append(_outputs(j))
        j.release()
    srv = eng.Server(ctx, max_in_flight=4, graph_passes=True, cobatch_tail=cobatch)
    try:
        for k, (cfg, cam, seed) in enumerate(_requests()):   # one at a time
            rid = srv.submit(cfg, cam, seed)
            j, _ = srv.wait(rid)
            _same(plain[k], _outputs(j), f"request {k} alone")
            j.release()
        for rep in range(2):  # everything at once: admitted together, overlapping on the device
            ids = [srv.submit(cfg, cam, seed) for cfg, cam, seed in _requests()]
            for k, rid in enumerate(ids):
                j, ms = srv.wait(rid)
                assert ms > 0
                _same(plain[k], _outputs(j), f"request {k} in the mix (round {rep})")
                j.release()
        st = srv.stats()
        assert st["completed"] == 15 and st["failed"] == 0
        assert st["max_in_flight_seen"] >= 2
        assert st["latency_p99_ms"] >= st["latency_p50_ms"] > 0
        n_statmc = sum(1 for cfg, _, _ in _requests() if cfg.use_statmc)
        if cobatch:   # every StatMC job's tail ran in a batch; requests without StatMC have none
            print(f"tail batches {st['tail_batches']}, jobs {st['tail_jobs']}")
            assert st["tail_jobs"] == 3 * n_statmc
            assert 1 <= st["tail_batches"] <= st["tail_jobs"]
        else:
            assert st["tail_batches"] == 0 and st["tail_jobs"] == 0
    finally:
        srv.close()
        ctx.close()


def test_render_finish_batch_identical_to_whole_jobs():
    """Passes on each job's stream, then one tail batch: uneven sizes, radii,
    adaptive-pass counts, a lens, a job without RPF capture (it takes the
    per-image tail) and a procedural-scene job; every output equals the whole
    pipeline's. A job without a pending tail is refused."""
    ctx = eng.Context(0)
    try:
        pc = eng.procedural_config(seed=9, subdivision_level=2, n_spheres=12)
        ctx.procedural_scene("proc-tail", pc)
        reqs = []
        for k, (w, h, spp, passes, ap, radius, shrink) in enumerate(
                [(64, 48, 4, 3, 0.0, 10, 1.0), (96, 64, 8, 2, 0.14, 5, 1.0), (40, 40, 2, 0, 0.0, 7, 1.0),
                 (56, 72, 4, 1, 0.0, 3, 0.0)]):
            rc = ob.config_c0(w, h)
            rc.spp, rc.adaptive_passes, rc.aperture, rc.focus_distance = spp, passes, ap, 1.05
            rc.color_window_radius, rc.rpf_shrinkage_scale = radius, shrink
            reqs.append(("cornell_box", engine_config(rc), engine_camera(rc), 500 + k))
        cfg = eng.default_config(image_width=72, image_height=40, samples_per_pixel=4, use_statmc=1,
                                 adaptive_passes=2, adaptive_spp=4, max_depth=8)
        reqs.append(("proc-tail", cfg, eng.make_camera((0.0, 0.6, 3.4), (0.0, -0.17, -1.0), (0.0, 1.0, 0.0), 40.0,
                                                      3.4, 0.06), 77))
        whole = []
        for scene, cfg, cam, seed in reqs:
            j = eng.Job(ctx, cfg, cam, seed, scene=scene, submit=True)
            j.wait()
            whole.append(_outputs(j))
            j.release()
        for rep in range(2):      # the second batch reuses the tuned StatMC shape
            jobs = [eng.Job(ctx, cfg, cam, seed, scene=scene, submit="passes") for scene, cfg, cam, seed in reqs]
            ms = eng.render_finish_batch(ctx, jobs)
            assert ms > 0
            for k, j in enumerate(jobs):
                _same(whole[k], _outputs(j), f"request {k} tail in a batch (round {rep})")
            with pytest.raises(eng.EngineError, match="no pending denoise tail"):
                eng.render_finish_batch(ctx, jobs[:1])
            for j in jobs:
                j.release()
    finally:
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


def test_slab_cache_takes_the_smallest_fitting_slab_within_the_held_limit():
    """A job's buffers are one slab. A released slab serves a later job that
    needs at most as much and at least two thirds of it (the job then holds
    the larger slab); with a held limit at the job's own size it takes only a
    slab of exactly its size. Beyond the cache cap the oldest slabs are freed."""
    ctx = eng.Context(0)
    cam = eng.make_camera()
    try:
        big = eng.default_config(image_width=96, image_height=64, samples_per_pixel=24, use_statmc=1)
        small = eng.default_config(image_width=96, image_height=64, samples_per_pixel=16, use_statmc=1)
        nb, ns = eng.job_bytes(big), eng.job_bytes(small)
        assert ns < nb <= ns + ns // 2, (ns, nb)
        assert ctx.set_buffer_cache(0) == 0
        assert ctx.set_buffer_cache(1 << 34) == 0
        j = eng.Job(ctx, big, cam, 1)
        j.release()
        assert ctx.set_buffer_cache(1 << 34) == nb
        j = eng.Job(ctx, small, cam, 2)                 # takes the larger kept slab
        assert ctx.memory()["in_use"] == nb
        assert ctx.set_buffer_cache(1 << 34) == 0
        j.release()
        ctx.set_held_limit(ns)                          # no room above its own size: a fresh slab
        j = eng.Job(ctx, small, cam, 3)
        assert ctx.memory()["in_use"] == ns
        assert ctx.set_buffer_cache(1 << 34) == nb
        j.release()
        ctx.set_held_limit(0)
        assert ctx.set_buffer_cache(1 << 34) == nb + ns
        assert ctx.set_buffer_cache(nb) == ns           # over the cap: the oldest (nb) goes first
        assert ctx.set_buffer_cache(0) == 0
    finally:
        ctx.close()
