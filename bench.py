#!/usr/bin/env python3
"""Benchmark of the path-trace -> RPF grid -> StatMC -> adaptive-pass job on one MI355X.

One step is one whole render+denoise job of the full-size plain workload:
built-in cornell_box, 1024x1024, 16 base spp + 3 adaptive passes of 4 spp,
StatMC with colour window radius 10, 3 variance-smoothing iterations and a
thin lens (aperture 0.14, focus 1.05), seed 1234, FP32 sample records.

Prints one JSON line. `value` is whole-job throughput in Msamples/s (path
samples traced per second of job time, all device state resident), one job
after another; the line also carries denoised MP/s, the same jobs through the
serving loop (several in flight: throughput, p50/p99 latency and image
checksums against the one-by-one images), the dominant kernel's roofline
(HIP-event timed inside the engine on the job's stream) and the CPU oracle
timed on a bounded sample of the same workload on this host.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(REPO, "rpd-statmc-denoiser_amd"))

import rpd_engine as eng  # noqa: E402

HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3
# FP32 operations of one watertight ray/triangle test (core/triangle.h:35-104),
# counting each +, -, *, / once: 9 vertex offsets, 6 shears (mul+sub = 12),
# U/V/W (6 mul + 3 sub), det (2 add), T (6 mul + 2 add), reciprocal and t (2),
# range tests (2).
TRI_TEST_FLOPS = 9 + 12 + 9 + 2 + 8 + 2 + 2
# One ray/sphere test (core/sphere.h:19-54): oc (3), a, b, c (15), disc (3),
# sqrt and root (4), range tests (2).
SPHERE_TEST_FLOPS = 3 + 15 + 3 + 4 + 2
TRAFFIC_FILE = os.path.join(REPO, "profiles", "k_trace_cornell1024_c1_traffic.md")

WORKLOAD = dict(width=1024, height=1024, spp=16, adaptive_spp=4, adaptive_passes=3, color_window_radius=10,
                var_iterations=3, aperture=0.14, focus_distance=1.05, seed=1234)


def job_config(w, h):
    return eng.default_config(
        image_width=w, image_height=h, samples_per_pixel=WORKLOAD["spp"], use_statmc=1,
        color_window_radius=WORKLOAD["color_window_radius"], var_iterations=WORKLOAD["var_iterations"],
        adaptive_spp=WORKLOAD["adaptive_spp"], adaptive_passes=WORKLOAD["adaptive_passes"])


def job_camera():
    return eng.make_camera(fov=38.6, focus_distance=WORKLOAD["focus_distance"], aperture=WORKLOAD["aperture"])


def trace_traffic_bytes():
    """HBM bytes per trace launch from the committed rocprofv3 FETCH_SIZE /
    WRITE_SIZE summary (copy-kernel calibrated), or None if it is absent."""
    try:
        with open(TRAFFIC_FILE) as f:
            for line in f:
                if line.startswith("traffic_bytes_per_launch:"):
                    return float(line.split(":", 1)[1])
    except OSError:
        return None
    return None


def check_sample_counter(job):
    """Timed steps count samples with the job's device counter (every traced path sample), which
    equals the per-pixel counts of finite samples unless some samples were non-finite. Checked
    once on a warm-up job, outside the timing."""
    finite = int(job.get("sample_count").sum())
    traced = job.counters()["samples"]
    if not finite <= traced <= finite * 1.001:
        print(f"note: {traced} traced samples against {finite} finite ones", file=sys.stderr)


def calibrate_copy(nbytes=1 << 30):
    """A plain device copy of known size, for scaling the FETCH_SIZE and
    WRITE_SIZE counters to bytes (run under rocprofv3 --pmc)."""
    import torch
    a = torch.empty(nbytes // 4, dtype=torch.float32, device="cuda").fill_(1.0)
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"calibration_copy_bytes": nbytes, "launches": 3}))


def cpu_baseline(seconds_hint=20.0):
    """The CPU oracle (a restatement of the reference's OpenMP path, including
    its overwritten in-loop StatMC calls) on one whole job of the same workload."""
    sys.path.insert(0, os.path.join(REPO, "tests"))
    import oracle_bind as ob
    threads = len(os.sched_getaffinity(0))      # every host core this process may run on (nproc)
    rc = ob.RenderConfig(width=WORKLOAD["width"], height=WORKLOAD["height"], spp=WORKLOAD["spp"], adaptive_spp=WORKLOAD["adaptive_spp"],
                         adaptive_passes=WORKLOAD["adaptive_passes"], color_window_radius=WORKLOAD["color_window_radius"],
                         var_iterations=WORKLOAD["var_iterations"], aperture=WORKLOAD["aperture"],
                         focus_distance=WORKLOAD["focus_distance"])
    t0 = time.perf_counter()
    j = ob.OracleJob(rc, WORKLOAD["seed"], nthreads=threads, skip_dead=False)
    j.render()
    dt = time.perf_counter() - t0
    samples = int(j.get("sample_count").sum())
    return {"value": samples / dt / 1e6, "unit": "Msamples/s", "cores": threads, "kind": "port",
            "sample": f"oracle, one whole job of the same workload ({samples} samples, {dt:.1f} s, "
                      f"{threads} OpenMP threads = all {threads} host cores from nproc)"}


REFERENCE_SPP = (8, 128)   # jobs x spp of the high-spp reference image (1024 spp)


def c1_quality(ctx, cfg, cam, w, h):
    """relMSE and NRMSE (shared/tools/eval_tools.cpp:146) of one job's final
    image and of its raw mean, against a 1024-spp uniform render of the same
    scene with other seeds (outside every timed region)."""
    ref = np.zeros((w * h, 3), np.float64)
    jobs, spp = REFERENCE_SPP
    rcfg = eng.default_config(image_width=w, image_height=h, samples_per_pixel=spp, use_statmc=0)
    for k in range(jobs):
        r = eng.Job(ctx, rcfg, cam, 7000 + k, submit=True)
        r.wait()
        ref += r.get("radiance")
        r.release()
    ref = (ref / jobs).astype(np.float32)
    j = eng.Job(ctx, cfg, cam, WORKLOAD["seed"], submit=True)
    j.wait()
    den, raw = j.metrics("radiance", ref), j.metrics("raw_radiance", ref)
    j.release()
    return {"reference": f"uniform render, {jobs * spp} spp, seeds 7000..{7000 + jobs - 1}",
            "relmse_denoised": den["relmse"], "relmse_raw": raw["relmse"],
            "nrmse_denoised": den["nrmse"], "nrmse_raw": raw["nrmse"]}


C2 = dict(width=1920, height=1080, spp=8, adaptive_spp=8, adaptive_passes=4, max_depth=8, aperture=0.06,
          focus_distance=3.4, fov=40.0, cam_pos=(0.0, 0.6, 3.4), cam_dir=(0.0, -0.17, -1.0), seed=1234)


def new_context(args):
    """The engine context with the trace and gather settings asked for on the command line."""
    ctx = eng.Context(0)
    if getattr(args, "trace_wavefront", None) is not None:
        ctx.set_trace_wavefront(args.trace_wavefront)
    if getattr(args, "gather_overlap", None) is not None:
        ctx.set_gather_overlap(args.gather_overlap)
    if getattr(args, "statmc_pairs", None) is not None:
        ctx.set_statmc_pairs(args.statmc_pairs)
    return ctx


def run_c2(args):
    """The procedural scene (seed 7: 1,310,720-triangle displaced icosphere, 256 GGX
    spheres, 4 area lights, constant environment) at 1920x1080, 8 spp + 4 adaptive
    passes of 8 spp, max depth 8 with Russian roulette, thin lens: one job per step."""
    ctx = new_context(args)
    pc = eng.procedural_config()
    t0 = time.perf_counter()
    ctx.procedural_scene("c2", pc)
    build_s = time.perf_counter() - t0
    cfg = eng.default_config(image_width=C2["width"], image_height=C2["height"], samples_per_pixel=C2["spp"],
                             use_statmc=1, adaptive_spp=C2["adaptive_spp"], adaptive_passes=C2["adaptive_passes"],
                             max_depth=C2["max_depth"])
    cam = eng.make_camera(C2["cam_pos"], C2["cam_dir"], (0.0, 1.0, 0.0), C2["fov"], C2["focus_distance"],
                          C2["aperture"])
    for _ in range(args.warmup):
        j = eng.Job(ctx, cfg, cam, C2["seed"], scene="c2", submit=True)
        j.wait()
        check_sample_counter(j)
        j.release()
    samples = queries = 0
    stage = np.zeros(11)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        j = eng.Job(ctx, cfg, cam, C2["seed"], scene="c2", submit=True)
        j.wait()
        stage += j.stage_ms()
        cnt = j.counters()
        samples += cnt["samples"]
        queries += cnt["queries"]
        j.release()
    wall = time.perf_counter() - t0
    n_px = C2["width"] * C2["height"]
    denoise_ms = (stage[2] + stage[3] + stage[4] + stage[5]) / args.steps
    result = {
        "metric": "Msamples/s & denoised MP/s; p50/p99 job latency vs plain at equal relMSE",
        "value": samples / wall / 1e6, "unit": "Msamples/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": wall * 1e3 / args.steps, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "fp32", "data": "synthetic: procedural scene generated on the host, rendered on device",
        "config": {"workload": "procedural scene seed 7 (1,310,720 triangles, 256 GGX spheres, 4 area lights, "
                               "environment), 1920x1080, 8 spp + 4 adaptive passes x 8 spp, max depth 8, thin lens",
                   "parallelism": "1 GPU"},
        "denoised_mp_per_s": n_px / (denoise_ms / 1e3) / 1e6 if denoise_ms > 0 else None,
        "samples_per_job": samples // args.steps,
        "ray_queries_per_s": queries / (stage[7] / 1e3) if stage[7] > 0 else None,
        "stage_ms": {"base_render": stage[0] / args.steps, "adaptive_render": stage[1] / args.steps,
                     "statmc": stage[3] / args.steps, "trace_kernel": stage[7] / args.steps,
                     "grid_gather": stage[10] / args.steps},
        "scene_build_s": build_s,
    }
    print(json.dumps(result))
    ctx.close()


C3 = dict(images=32, side=(256, 2048), spp=(4, 64), seed=7, budget_gb=265.0)


def synth_truth(albedo, sc):
    """Expected radiance of a synthetic pixel: the lognormal mixture's mean
    exp(mu + sigma^2/2) * (1 - f + f * scale), tinted by the region albedo."""
    m = np.exp(sc.log_mean + 0.5 * sc.log_sigma ** 2) * (1 - sc.firefly_fraction + sc.firefly_fraction * sc.firefly_scale)
    a = albedo.astype(np.float64)
    return m * a / a.mean(axis=1, keepdims=True)


def relmse(x, ref):
    """mean over pixels and channels of (x - r)^2 / (r^2 + 1e-2)."""
    x = x.astype(np.float64)
    return float(np.mean((x - ref) ** 2 / (ref ** 2 + 1e-2)))


def run_c3(args):
    """Batched denoise of 32 synthetic per-sample buffers (side U{256..2048},
    spp U{4..64}): the plain per-image path (one rpd_job_denoise_records call
    per image, one image after another) against rpd_denoise_batch (every stage
    one launch over all images). Both are timed the same way, host wall clock
    around the calls (each returns with its work done), after a warm-up of
    each, over args.steps repetitions on freshly synthesised jobs; medians."""
    rng = np.random.default_rng(C3["seed"])
    sides = rng.integers(C3["side"][0], C3["side"][1] + 1, C3["images"])
    spps = rng.integers(C3["spp"][0], C3["spp"][1] + 1, C3["images"])
    ctx = new_context(args)
    cfg0 = eng.default_config(image_width=8, image_height=8, use_statmc=1)
    radius = cfg0.color_window_radius
    ctx.set_memory_budget(int(C3["budget_gb"] * 1e9))
    scs = [eng.synth_config(seed=1000 + k, samples_per_pixel=int(spps[k])) for k in range(C3["images"])]
    n_px = int((sides.astype(np.int64) ** 2).sum())
    n_samples = int((sides.astype(np.int64) ** 2 * spps).sum())

    def fresh_jobs():
        jobs = []
        for k in range(C3["images"]):
            side, spp = int(sides[k]), int(spps[k])
            cfg = eng.default_config(image_width=side, image_height=side, samples_per_pixel=spp, use_statmc=1)
            jobs.append(eng.Job(ctx, cfg, None, 0, denoiser_spp=spp))
            jobs[-1].synthesize_records(scs[k])
        return jobs

    def plain_run():
        jobs = fresh_jobs()
        t = time.perf_counter()
        for j in jobs:
            j.denoise_records()
        return (time.perf_counter() - t) * 1e3, jobs

    def batch_run():
        jobs = fresh_jobs()
        t = time.perf_counter()
        dev_ms = eng.denoise_batch(ctx, jobs)
        return (time.perf_counter() - t) * 1e3, dev_ms, jobs

    print(f"c3: {C3['images']} images, {n_px / 1e6:.1f} Mpx, {n_samples / 1e9:.2f} G samples", file=sys.stderr,
          flush=True)
    plain_ms, batch_ms, batch_dev_ms = [], [], []
    plain = None
    for rep in range(1 + args.steps):                 # repetition 0 is the warm-up of both paths
        p_ms, jobs = plain_run()
        if rep == 0:
            plain = [j.get("denoised") for j in jobs]
        else:
            plain_ms.append(p_ms)
        for j in jobs:
            j.release()
        b_ms, dev, jobs = batch_run()
        if rep > 0:
            batch_ms.append(b_ms)
            batch_dev_ms.append(dev)
        if rep < args.steps:
            for j in jobs:
                j.release()
        print(f"c3: repetition {rep}: plain {p_ms:.0f} ms, batched {b_ms:.0f} ms", file=sys.stderr, flush=True)
    identical = all(np.array_equal(a, j.get("denoised")) for a, j in zip(plain, jobs))
    err_den = err_in = 0.0
    for j, sc in zip(jobs, scs):
        truth = synth_truth(j.get("albedo") / np.maximum(j.get("hit_count"), 1)[:, None], sc)
        wgt = j.w * j.h_px / n_px
        err_den += wgt * relmse(j.get("denoised"), truth)
        err_in += wgt * relmse(j.get("stats")[:, 3:6], truth)
    for j in jobs:
        j.release()
    batch = float(np.median(batch_ms))
    plain_med = float(np.median(plain_ms))
    result = {
        "metric": "denoised MP/s, batched denoise of synthetic per-sample buffers",
        "value": n_px / (batch / 1e3) / 1e6, "unit": "MP/s", "n_gpus": 1, "steps": args.steps, "warmup": 1,
        "ms_per_step": batch, "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "fp32+fp64",
        "data": "synthetic: counter-based per-sample records generated on device",
        "config": {"workload": f"{C3['images']} images, side U{{{C3['side'][0]}..{C3['side'][1]}}}, "
                               f"spp U{{{C3['spp'][0]}..{C3['spp'][1]}}}, seed {C3['seed']}, StatMC r{radius}",
                   "megapixels": n_px / 1e6, "gsamples": n_samples / 1e9},
        "timing": "host wall clock around the calls for both paths, median of steps after one warm-up",
        "plain_ms": plain_med, "plain_mp_per_s": n_px / (plain_med / 1e3) / 1e6,
        "batched_device_ms": float(np.median(batch_dev_ms)), "speedup_vs_plain": plain_med / batch,
        "identical_to_plain": identical,
        "relmse": {"denoised": err_den, "input_mean": err_in},
        "filter_tile": ({"radius": radius, "form": "pair and sum passes", "tile": [16, 16],
                         "chosen_by": "fixed: the batched pair form walks 16x16 tiles"}
                        if getattr(args, "statmc_pairs", None) != 0 else
                        {"radius": radius, "form": "one pass", "tile": list(ctx.statmc_tile(radius, n_px)),
                         "chosen_by": "autotuner (every tile shape timed on the first batch)"}),
    }
    print(json.dumps(result))
    ctx.close()


C4 = dict(jobs=64, rate=20.0, sides=(512, 1024, 1536, 2048), spp=(4, 8, 16, 32), passes=(0, 4),
          procedural_fraction=1.0 / 3.0, budget_gb=200.0, max_in_flight=8, seed=11)


def c4_requests():
    """The fixed request stream: Poisson arrival times and per-job shapes drawn
    from one seed; a third of the jobs render the procedural scene (C2 camera,
    max depth 8), the others cornell_box (C0/C1 camera, aperture 0.14)."""
    rng = np.random.default_rng(C4["seed"])
    t = np.cumsum(rng.exponential(1.0 / C4["rate"], C4["jobs"]))
    reqs = []
    for k in range(C4["jobs"]):
        side = int(rng.choice(C4["sides"]))
        spp = int(rng.choice(C4["spp"]))
        passes = int(rng.integers(C4["passes"][0], C4["passes"][1] + 1))
        procedural = bool(rng.random() < C4["procedural_fraction"])
        kw = dict(image_width=side, image_height=side, samples_per_pixel=spp, use_statmc=1, adaptive_spp=4,
                  adaptive_passes=passes)
        if procedural:
            cfg = eng.default_config(max_depth=C2["max_depth"], **kw)
            cam = eng.make_camera(C2["cam_pos"], C2["cam_dir"], (0.0, 1.0, 0.0), C2["fov"], C2["focus_distance"],
                                  C2["aperture"])
        else:
            cfg = eng.default_config(color_window_radius=10, var_iterations=3, **kw)
            cam = job_camera()
        reqs.append((float(t[k]), "c2" if procedural else "cornell_box", cfg, cam, 5000 + k))
    return reqs


def c4_stream(ctx, reqs, rate_scale, max_in_flight, cobatch_tail=False):
    """The request stream through the serving loop, arrivals at their Poisson
    times divided by rate_scale (1 = the stream's own rate); returns wall
    seconds, samples, the loop's statistics and every job's image checksum.
    cobatch_tail: the denoise tails of jobs whose passes finish together run
    as one batch."""
    import threading
    srv = eng.Server(ctx, max_in_flight=max_in_flight, memory_budget=int(C4["budget_gb"] * 1e9), graph_passes=True,
                     cobatch_tail=cobatch_tail)
    ids = [None] * len(reqs)
    samples = [0] * len(reqs)
    sums = [0.0] * len(reqs)

    def submitter(t0):
        for k, (at, scene, cfg, cam, seed) in enumerate(reqs):
            delay = t0 + at / rate_scale - time.perf_counter()
            if delay > 0:
                time.sleep(delay)
            ids[k] = srv.submit(cfg, cam, seed, scene=scene)

    t0 = time.perf_counter()
    th = threading.Thread(target=submitter, args=(t0,))
    th.start()
    for k in range(len(reqs)):
        while ids[k] is None:
            time.sleep(0.001)
        j, _ = srv.wait(ids[k])
        samples[k] = int(j.get("sample_count").sum())
        sums[k] = float(np.asarray(j.get("radiance"), np.float64).sum())
        j.release()
        if (k + 1) % 16 == 0:
            print(f"c4: {k + 1}/{len(reqs)} jobs done at {time.perf_counter() - t0:.1f} s", file=sys.stderr, flush=True)
    wall = time.perf_counter() - t0
    th.join()
    st = srv.stats()
    srv.close()
    return wall, sum(samples), st, sums


SERVED_C1 = dict(jobs=16, max_in_flight=6, sustainable_fraction=0.8, seed=21)


def served_c1(ctx, cfg, cam, plain_ms, plain_sum, plain_samples):
    """The full-size job through the serving loop (rpd_server): 16 requests of
    the same job (seed 1234), first all at once (the saturated throughput, 6 in
    flight), then as Poisson arrivals at 0.8 of that throughput (the latency of
    a loop that keeps up). Every served image must equal the one-by-one image
    (radiance checksum), so relMSE is the same in both modes."""
    import threading

    def stream(arrivals):
        srv = eng.Server(ctx, max_in_flight=SERVED_C1["max_in_flight"], graph_passes=False)
        ids = [None] * len(arrivals)

        def submitter(t0):
            for k, at in enumerate(arrivals):
                delay = t0 + at - time.perf_counter()
                if delay > 0:
                    time.sleep(delay)
                ids[k] = srv.submit(cfg, cam, WORKLOAD["seed"])

        t0 = time.perf_counter()
        th = threading.Thread(target=submitter, args=(t0,))
        th.start()
        sums, samples = [], 0
        for k in range(len(arrivals)):
            while ids[k] is None:
                time.sleep(0.0005)
            j, _ = srv.wait(ids[k])
            sums.append(float(np.asarray(j.get("radiance"), np.float64).sum()))
            samples += int(j.get("sample_count").sum())
            j.release()
        wall = time.perf_counter() - t0
        th.join()
        st = srv.stats()
        srv.close()
        return wall, samples, sums, st

    n = SERVED_C1["jobs"]
    sat_wall, sat_samples, sat_sums, sat = stream([0.0] * n)
    rate = n / sat_wall
    rng = np.random.default_rng(SERVED_C1["seed"])
    arrivals = list(np.cumsum(rng.exponential(1.0 / (SERVED_C1["sustainable_fraction"] * rate), n)))
    sus_wall, sus_samples, sus_sums, sus = stream(arrivals)
    lat = lambda st: {"p50": st["latency_p50_ms"], "p99": st["latency_p99_ms"], "max": st["latency_max_ms"],
                      "queue_p99": st["queue_p99_ms"], "max_in_flight_seen": st["max_in_flight_seen"]}
    return {
        "jobs": n, "max_in_flight": SERVED_C1["max_in_flight"],
        "jobs_per_s": rate, "msamples_per_s": sat_samples / sat_wall / 1e6, "ms_per_job": sat_wall * 1e3 / n,
        "over_plain": plain_ms / (sat_wall * 1e3 / n),
        "latency_ms_all_at_once": lat(sat),
        "sustainable": {"arrival_jobs_per_s": SERVED_C1["sustainable_fraction"] * rate, "latency_ms": lat(sus)},
        "plain_latency_ms": plain_ms,
        "images_equal_to_plain": all(v == plain_sum for v in sat_sums + sus_sums),
        "samples_equal_to_plain": sat_samples == n * plain_samples and sus_samples == n * plain_samples,
    }


def run_c4(args):
    """C4: one fixed request stream four ways, on one context:
      * plain: the per-image GPU path, one whole job after another
        (rpd_job_submit + wait, no overlap), arrival times ignored;
      * served at the stream's rate (20 jobs/s: more than the device serves,
        so this is the saturated throughput and its queueing latency);
      * the same with the denoise tails of jobs whose passes finish together
        batched into shared launches;
      * served at a sustainable rate, 0.8 of the plain path's job rate.
    Served jobs equal the plain ones bit for bit (a job's result does not
    depend on what else runs), so the images, and their relMSE, are the same;
    the checksums of every image are compared."""
    ctx = new_context(args)
    ctx.procedural_scene("c2", eng.procedural_config())
    reqs = c4_requests()
    # warm-up: one small job per scene (code objects, graphs, allocator, StatMC tile tuning)
    for scene in ("cornell_box", "c2"):
        for r in (10, 7):
            cfg = eng.default_config(image_width=256, image_height=256, samples_per_pixel=4, use_statmc=1,
                                     adaptive_spp=4, adaptive_passes=1, color_window_radius=r,
                                     max_depth=8 if scene == "c2" else 0)
            j = eng.Job(ctx, cfg, job_camera(), 1, scene=scene, submit=True)
            j.wait()
            j.release()
    # plain per-image path
    t0 = time.perf_counter()
    plain_samples = 0
    plain_sums = []
    for k, (_, scene, cfg, cam, seed) in enumerate(reqs):
        j = eng.Job(ctx, cfg, cam, seed, scene=scene, submit=True)
        j.wait()
        plain_samples += int(j.get("sample_count").sum())
        plain_sums.append(float(np.asarray(j.get("radiance"), np.float64).sum()))
        j.release()
        if (k + 1) % 16 == 0:
            print(f"c4 plain: {k + 1}/{len(reqs)} jobs done at {time.perf_counter() - t0:.1f} s", file=sys.stderr,
                  flush=True)
    plain_wall = time.perf_counter() - t0
    plain_rate = len(reqs) / plain_wall
    # served, at the stream's rate and at a sustainable one
    sat_wall, sat_samples, sat, sat_sums = c4_stream(ctx, reqs, 1.0, C4["max_in_flight"])
    cob_wall, cob_samples, cob, cob_sums = c4_stream(ctx, reqs, 1.0, C4["max_in_flight"], cobatch_tail=True)
    arrival_rate = len(reqs) / reqs[-1][0]
    sus_scale = 0.8 * plain_rate / arrival_rate
    sus_wall, sus_samples, sus, sus_sums = c4_stream(ctx, reqs, sus_scale, C4["max_in_flight"])
    same = sat_sums == plain_sums and sus_sums == plain_sums and cob_sums == plain_sums
    n_proc = sum(1 for r in reqs if r[1] == "c2")

    def lat(st):
        return {"p50": st["latency_p50_ms"], "p99": st["latency_p99_ms"], "mean": st["latency_mean_ms"],
                "max": st["latency_max_ms"], "queue_p50": st["queue_p50_ms"], "queue_p99": st["queue_p99_ms"],
                "max_in_flight_seen": st["max_in_flight_seen"], "peak_gb": st["peak_bytes"] / 1e9,
                "admitted_over_budget": st["over_budget"],
                "failed": st["failed"]}
    result = {
        "metric": "Msamples/s & denoised MP/s; p50/p99 job latency vs plain at equal relMSE",
        "value": sat_samples / sat_wall / 1e6, "unit": "Msamples/s", "n_gpus": 1, "steps": len(reqs), "warmup": 4,
        "ms_per_step": sat_wall * 1e3 / len(reqs), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "fp32", "data": "synthetic: built-in cornell_box and procedural scenes, rendered on device",
        "config": {"workload": f"{len(reqs)} jobs, Poisson arrivals at {C4['rate']:g}/s (seed {C4['seed']}), side "
                               f"U{set(C4['sides'])}, base spp U{set(C4['spp'])}, adaptive passes "
                               f"U{{{C4['passes'][0]}..{C4['passes'][1]}}} x 4 spp, {n_proc} procedural-scene jobs, "
                               f"HBM budget {C4['budget_gb']:g} GB, {C4['max_in_flight']} in flight",
                   "parallelism": "1 GPU, concurrent jobs on separate streams"},
        "plain": {"jobs_per_s": plain_rate, "msamples_per_s": plain_samples / plain_wall / 1e6,
                  "wall_s": plain_wall},
        "served_20_per_s": {"jobs_per_s": len(reqs) / sat_wall, "msamples_per_s": sat_samples / sat_wall / 1e6,
                            "wall_s": sat_wall, "latency_ms": lat(sat)},
        "served_20_per_s_cobatched_tail": {
            "jobs_per_s": len(reqs) / cob_wall, "msamples_per_s": cob_samples / cob_wall / 1e6, "wall_s": cob_wall,
            "latency_ms": lat(cob), "tail_batches": cob["tail_batches"],
            "jobs_per_tail_batch": cob["tail_jobs"] / max(1, cob["tail_batches"])},
        "served_sustainable": {"arrival_jobs_per_s": arrival_rate * sus_scale,
                               "jobs_per_s": len(reqs) / sus_wall, "msamples_per_s": sus_samples / sus_wall / 1e6,
                               "wall_s": sus_wall, "latency_ms": lat(sus)},
        "served_over_plain": (len(reqs) / sat_wall) / plain_rate,
        "cobatched_over_plain": (len(reqs) / cob_wall) / plain_rate,
        "images_equal_to_plain": same,
    }
    print(json.dumps(result))
    ctx.close()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-quality", action="store_true", help="skip the relMSE against a high-spp reference")
    ap.add_argument("--no-served", action="store_true", help="skip the same jobs through the serving loop")
    ap.add_argument("--size", type=int, default=WORKLOAD["width"])
    ap.add_argument("--trace-wavefront", type=float, default=None,
                    help="wavefront trace rounds per sample for flat scenes (0: path-regeneration kernel only; "
                         "default: the engine's default)")
    ap.add_argument("--gather-overlap", type=int, choices=[0, 1], default=None,
                    help="grid gather of a pass beside the next pass's trace (1) or in line (0); "
                         "default: the engine's default")
    ap.add_argument("--statmc-pairs", type=int, choices=[0, 1], default=None,
                    help="StatMC window as the symmetric pair and sum passes (1) or one pass (0); "
                         "default: the engine's default")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="run a 1 GiB device copy (counter calibration) and exit")
    ap.add_argument("--workload", default="c1", choices=["c1", "c2", "c3", "c4"],
                    help="c1: full-size cornell_box render+denoise job (default); c2: the 1.3M-triangle "
                         "procedural scene at 1920x1080; c3: batched denoise of 32 synthetic per-sample buffers; "
                         "c4: a Poisson stream of mixed render+denoise jobs through the serving loop")
    args = ap.parse_args()
    if args.calibrate_copy:
        calibrate_copy()
        return
    if args.workload == "c2":
        if args.gpus != 1:
            print("bench.py measures one GPU", file=sys.stderr)
            sys.exit(2)
        run_c2(args)
        return
    if args.workload in ("c3", "c4"):
        if args.gpus != 1:
            print("bench.py measures one GPU", file=sys.stderr)
            sys.exit(2)
        (run_c3 if args.workload == "c3" else run_c4)(args)
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU (asked for --gpus {args.gpus}): this workload runs on a single "
              "MI355X; sharding it over several GPUs is not implemented", file=sys.stderr)
        sys.exit(2)
    w = h = args.size
    ctx = new_context(args)
    cfg, cam = job_config(w, h), job_camera()
    for _ in range(args.warmup):
        j = eng.Job(ctx, cfg, cam, WORKLOAD["seed"], submit=True)
        j.wait()
        check_sample_counter(j)
        j.release()
    samples = 0
    queries = 0
    stage = np.zeros(11)
    t0 = time.perf_counter()
    for k in range(args.steps):
        j = eng.Job(ctx, cfg, cam, WORKLOAD["seed"], submit=True)
        j.wait()
        stage += j.stage_ms()
        cnt = j.counters()
        samples += cnt["samples"]
        queries += cnt["queries"]
        if k < args.steps - 1:
            j.release()
    wall = time.perf_counter() - t0
    # the last job's image, read after the timed region (for the served comparison)
    plain_sum = float(np.asarray(j.get("radiance"), np.float64).sum())
    plain_samples = int(j.get("sample_count").sum())
    j.release()
    ms_per_step = wall * 1e3 / args.steps
    n_px = w * h
    # denoised MP/s over the StatMC-side stages (sensitivity, StatMC, variance, allocation)
    denoise_ms = (stage[2] + stage[3] + stage[4] + stage[5]) / args.steps
    trace_ms_total = stage[7] / args.steps
    result = {
        "metric": "Msamples/s & denoised MP/s; p50/p99 job latency vs plain at equal relMSE",
        "value": samples / wall / 1e6,
        "unit": "Msamples/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms_per_step,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "fp32",
        "data": "synthetic: built-in cornell_box scene, rendered on device",
        "config": {"workload": f"cornell_box {w}x{h}, 16 base spp + 3 adaptive passes x 4 spp, StatMC r10, "
                               f"3 variance iterations, aperture 0.14, seed 1234",
                   "parallelism": "1 GPU"},
        "denoised_mp_per_s": n_px / (denoise_ms / 1e3) / 1e6 if denoise_ms > 0 else None,
        "jobs_per_s": args.steps / wall,
        "stage_ms": {"base_render": stage[0] / args.steps, "adaptive_render": stage[1] / args.steps,
                     "sensitivity": stage[2] / args.steps, "statmc": stage[3] / args.steps,
                     "variance": stage[4] / args.steps, "allocation": stage[5] / args.steps,
                     "trace_kernel": trace_ms_total, "grid_gather": stage[10] / args.steps},
        "samples_per_job": samples // args.steps,
        "statmc_tile": list(ctx.statmc_tile(WORKLOAD["color_window_radius"], n_px)),
    }
    # roofline of the dominant kernel, the trace kernel: FP32 VALU-bound.
    # Algorithmic work = ray queries x (triangles x TRI_TEST_FLOPS + spheres x
    # SPHERE_TEST_FLOPS), over the HIP-event time of its launches.
    launches = stage[8]
    if launches > 0 and stage[7] > 0:
        flops = queries * (cnt["triangles"] * TRI_TEST_FLOPS + cnt["spheres"] * SPHERE_TEST_FLOPS)
        achieved = flops / (stage[7] / 1e3) / 1e12
        avg_launch_s = stage[7] / launches / 1e3
        traffic = trace_traffic_bytes()
        result["roofline"] = {"bound": "valu", "achieved": achieved, "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                              "frac": achieved / FP32_PEAK_TFLOPS, "traffic": traffic,
                              "kernel": "k_trace (FP32 vector ALU, no matrix-core work; FP32 vector peak)",
                              "avg_launch_ms": stage[7] / launches,
                              "ray_queries_per_job": queries // args.steps,
                              "ray_queries_per_s": queries / (stage[7] / 1e3)}
        if traffic:
            # the same kernel against the HBM roof: counter bytes per launch over its launch time
            result["roofline"]["hbm_gbs"] = traffic / avg_launch_s / 1e9
            result["roofline"]["hbm_frac"] = traffic / avg_launch_s / 1e9 / HBM_PEAK_GBS
    if not args.no_served:
        result["served"] = served_c1(ctx, cfg, cam, ms_per_step, plain_sum, plain_samples)
    if not args.no_quality:
        result["quality"] = c1_quality(ctx, cfg, cam, w, h)
    if not args.no_cpu_baseline:
        result["cpu_baseline"] = cpu_baseline()
    print(json.dumps(result))
    ctx.close()


if __name__ == "__main__":
    main()
