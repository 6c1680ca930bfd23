"""Pins of the CPU oracle (CPU only).

The oracle restates the reference renderer; these tests tie it to what fixes
the reference's behaviour independently of the restatement:
  * the reference's own known-answer tests (path_tracer/tests/stats_test.cpp:11-76);
  * the C++ standard's mt19937 known answer and libstdc++'s float mapping;
  * textbook mathematics (Student-t quantiles, concentric disk/sphere maps,
    BRDF normalisation, explained-variance definitions by a two-pass method);
  * brute-force checks (BVH vs all-triangle intersection);
  * invariants the reference's scripts assert (paper_eval.sh:195-196).
"""
import ctypes as C

import numpy as np
import pytest
from scipy import stats as sps

import oracle_bind as ob
from oracle_bind import ptr

L = ob.lib()


# ---- sampler ---------------------------------------------------------------------------
def test_mt19937_known_answer():
    # C++ [rand.predef]: the 10000th output of a default-constructed mt19937 is 4123659995
    assert L.orc_mt_u32(5489, 10000) == 4123659995


def test_draw_is_float_of_u32_over_2_32():
    # libstdc++ generate_canonical<float,24> with a 32-bit engine: float(u) / 2^32
    out = np.zeros(8, np.float32)
    L.orc_sampler_draws(1234, 8, ptr(out))
    for k in range(8):
        u = L.orc_mt_u32(1234, k + 1)
        assert out[k] == np.float32(np.float32(u) / np.float32(2.0 ** 32))
    assert (out < 1.0).all() and (out >= 0.0).all()


def test_pair_order_is_second_draw_first():
    # Vec2f(next1d(), next1d()) under GCC evaluates the second argument first
    d = np.zeros(4, np.float32)
    p = np.zeros(4, np.float32)
    L.orc_sampler_draws(77, 4, ptr(d))
    L.orc_sampler_pairs(77, 2, ptr(p))
    assert p[0] == d[1] and p[1] == d[0] and p[2] == d[3] and p[3] == d[2]


def test_seed_hash_is_bijective_and_unit_float_range():
    xs = np.arange(0, 1 << 16, dtype=np.uint32) * np.uint32(65537)
    hs = {L.orc_hash_u32(int(x)) for x in xs[:4096]}
    assert len(hs) == 4096                   # each xorshift-multiply step is invertible
    assert L.orc_hash_u32(0) == 0
    f = [L.orc_hash_to_unit_float(int(x)) for x in xs[:2048]]
    assert min(f) >= 0.0 and max(f) < 1.0
    # 24-bit resolution
    assert all(float(v) * (1 << 24) == int(float(v) * (1 << 24)) for v in f)


# ---- reference known-answer tests: path_tracer/tests/stats_test.cpp -----------------------
def test_welford_color_mean_known_answer():        # stats_test.cpp:12-17
    rgb = np.array([[1, 4, 9], [3, 16, 25]], np.float32)
    out = np.zeros(12, np.float32)
    L.orc_accumulate_samples(ptr(rgb), 2, ptr(out))
    assert out[2] == 2
    assert np.linalg.norm(out[3:6] - np.array([2, 10, 17])) < 1e-5
    assert out[6:9].min() > 0.0                      # sqrt_color_variance
    # sqrt-domain Welford equals the two-pass sample variance of sqrt(c)
    s = np.sqrt(rgb.astype(np.float64))
    np.testing.assert_allclose(out[6:9], s.var(axis=0, ddof=1), rtol=1e-5)
    np.testing.assert_allclose(out[9:12], s.mean(axis=0), rtol=1e-6)


def test_rpf_helpers_known_answers():                # stats_test.cpp:18-29
    assert abs(L.orc_relative_random_sensitivity(0.5, 0.5) - 0.5) < 1e-5
    assert L.orc_relative_random_sensitivity(0.0, 0.5) == 0.0
    lhs = np.array([1, 0, 0, 0, 0, 0], np.float32)
    shared = np.array([0.5, 0, 0, 0, 0, 0], np.float32)
    disjoint = np.array([0, 1, 0, 0, 0, 0], np.float32)
    assert L.orc_shared_reliability(ptr(lhs), ptr(shared)) == 0.5
    assert L.orc_shared_reliability(ptr(lhs), ptr(disjoint)) == 0.0
    assert L.orc_relax_compatibility(4.0, 1.0, 0.0) == 4.0
    assert L.orc_relax_compatibility(4.0, 1.0, 1.0) == 2.0
    assert L.orc_relax_compatibility(4.0, 0.0, 1.0) == 4.0
    assert L.orc_conservative_variance_update(1.0, 2.0, 1e-3) > 1.0
    assert L.orc_conservative_variance_update(2.0, 1.0, 1e-3) == 2.0


def test_mirror_sample_weight_is_albedo():          # stats_test.cpp:31-37
    alb = np.array([0.8, 0.7, 0.6], np.float32)
    n = np.array([0, 0, 1], np.float32)
    wo = np.array([0.6, 0, 0.8], np.float32)
    out = np.zeros(7, np.float32)
    L.orc_brdf_sample(2, ptr(alb), 0.0, ptr(wo), ptr(n), 0.0, 0.0, ptr(out))
    assert np.linalg.norm(out[4:7] - alb) < 1e-5
    np.testing.assert_allclose(out[0:3], [-0.6, 0, 0.8], atol=1e-6)   # reflection about n


def test_dependency2d_checkerboard_known_answer():  # stats_test.cpp:39-50
    u, v, y1, y2 = [], [], [], []
    for yy in range(40):
        for xx in range(40):
            u.append((xx + 0.5) / 40)
            v.append((yy + 0.5) / 40)
            y1.append(1.0 if ((xx // 10) + (yy // 10)) % 2 else 0.0)
            y2.append(1.0 if (xx + 3 * yy) % 2 else 0.0)
    u, v = np.array(u, np.float32), np.array(v, np.float32)
    assert L.orc_dep2d(ptr(u), ptr(v), ptr(np.array(y1, np.float32)), 1600) > 0.95
    assert L.orc_dep2d(ptr(u), ptr(v), ptr(np.array(y2, np.float32)), 1600) < 0.05


def _binned(nbins, bins, y, w=None):
    bins = np.asarray(bins, np.int32)
    y = np.asarray(y, np.float32)
    w = np.ones(len(bins), np.float32) if w is None else np.asarray(w, np.float32)
    out = np.zeros(2, np.float64)
    L.orc_binned(nbins, ptr(bins), ptr(y), ptr(w), len(bins), ptr(out))
    return out


def test_tile_light_known_answers():                 # stats_test.cpp:52-63
    # light_uv: 10 samples, 5 at u=0.1 with y=0 and 5 at u=0.9 with y=1 (bins 0 and 15)
    bins = [0] * 5 + [15] * 5
    y = [0.0] * 5 + [1.0] * 5
    sens, mass = _binned(16, bins, y)
    assert sens > 0.95
    assert abs(L.orc_confidence(mass, 10.0) - 0.5) < 1e-5
    sel_sens, sel_mass = _binned(16, [0] * 100, [float(i % 2) for i in range(100)])
    assert sel_sens == 0.0
    assert L.orc_confidence(sel_mass, 10.0) > 0.9


def test_binned2_known_answer():                     # stats_test.cpp:65-74
    bins = [b for b in range(2) for _ in range(100)]
    dep = [float(b) for b in bins]
    ind = [float(i % 2) for b in range(2) for i in range(100)]
    assert _binned(2, bins, dep)[0] > 0.95
    assert _binned(2, bins, ind)[0] < 0.05


def test_binned_sensitivity_two_pass_definition():
    """Weighted explained variance by group means (two-pass) vs the moment sums."""
    rng = np.random.default_rng(5)
    for trial in range(20):
        n = int(rng.integers(5, 60))
        bins = rng.integers(0, 16, n)
        y = rng.normal(size=n).astype(np.float32) + (bins % 3) * rng.uniform(0, 2)
        w = rng.uniform(0.1, 1.0, n).astype(np.float32)
        got, mass = _binned(16, bins, y, w)
        wd, yd = w.astype(np.float64), y.astype(np.float64)
        mean = (wd * yd).sum() / wd.sum()
        total = (wd * (yd - mean) ** 2).sum()
        between = 0.0
        for b in np.unique(bins):
            m = bins == b
            between += wd[m].sum() * ((wd[m] * yd[m]).sum() / wd[m].sum() - mean) ** 2
        neff = wd.sum() ** 2 / (wd ** 2).sum()
        k = len(np.unique(bins))
        assert abs(mass - neff) <= 1e-9 * neff
        if neff <= 2 or k < 2 or neff <= k or total <= 1e-18:
            assert got == 0.0
            continue
        eta2 = min(max(between / total, 0.0), 1.0)
        want = np.sqrt(max(0.0, 1 - (1 - eta2) * (neff - 1) / (neff - k)))
        assert abs(got - want) <= 1e-6


# ---- t critical values: main.cpp:188-231 against Student-t quantiles ------------------------
def test_t_critical_tables_match_student_t():
    for df in range(1, 31):
        assert abs(L.orc_t_critical(float(df), 0.05) - sps.t.ppf(0.975, df)) < 1.5e-3
        assert abs(L.orc_t_critical(float(df), 0.01) - sps.t.ppf(0.995, df)) < 1.5e-3
    # linear interpolation between table rows
    mid = L.orc_t_critical(2.5, 0.05)
    assert abs(mid - 0.5 * (4.303 + 3.182)) < 1e-5
    # beyond the table: normal quantile, and df < 1 clamps to 1
    assert abs(L.orc_t_critical(100.0, 0.05) - 1.96) < 1e-6
    assert L.orc_t_critical(0.2, 0.05) == L.orc_t_critical(1.0, 0.05)
    # other alphas: normal quantile step with the small-df inflation sqrt(df/(df-1))
    assert abs(L.orc_t_critical(100.0, 0.2) - 1.28) < 1e-6
    assert abs(L.orc_t_critical(10.0, 0.1) - 1.64 * np.sqrt(10 / 9)) < 1e-5


# ---- sampling maps and BRDFs --------------------------------------------------------------
def test_concentric_disk_map():
    out = np.zeros(2, np.float32)
    L.orc_sample_disk(0.5, 0.5, ptr(out))
    assert (out == 0).all()
    L.orc_sample_disk(1.0, 0.5, ptr(out))
    np.testing.assert_allclose(out, [1, 0], atol=1e-6)
    rng = np.random.default_rng(1)
    r = []
    for u in rng.random((4000, 2)).astype(np.float32):
        L.orc_sample_disk(float(u[0]), float(u[1]), ptr(out))
        r.append(np.hypot(*out))
    r = np.array(r)
    assert r.max() <= 1.0 + 1e-6
    # uniform on the disk: P(r <= t) = t^2
    assert sps.kstest(r ** 2, "uniform").pvalue > 1e-3


def test_uniform_sphere_map():
    out = np.zeros(3, np.float32)
    rng = np.random.default_rng(2)
    zs = []
    for u in rng.random((4000, 2)).astype(np.float32):
        L.orc_sample_sphere(float(u[0]), float(u[1]), ptr(out))
        assert abs(np.linalg.norm(out) - 1) < 1e-5
        zs.append(out[2])
    assert sps.kstest((np.array(zs) + 1) / 2, "uniform").pvalue > 1e-3


def test_lambertian_sampling_weight_is_albedo_and_pdf_is_cosine():
    alb = np.array([0.3, 0.5, 0.7], np.float32)
    n = np.array([0.0, 0.6, 0.8], np.float32)
    wo = np.array([0.0, 0.0, 1.0], np.float32)
    out = np.zeros(7, np.float32)
    rng = np.random.default_rng(3)
    for u in rng.random((200, 2)).astype(np.float32):
        L.orc_brdf_sample(0, ptr(alb), 0.5, ptr(wo), ptr(n), float(u[0]), float(u[1]), ptr(out))
        wi, pdf = out[0:3], out[3]
        assert abs(np.linalg.norm(wi) - 1) < 1e-5
        cos = float(np.dot(wi, n))
        assert abs(pdf - cos / np.pi) < 1e-5
        np.testing.assert_allclose(out[4:7], alb, rtol=1e-4, atol=1e-6)


@pytest.mark.parametrize("kind,rough", [(1, 0.3), (1, 0.8), (3, 0.4)])
def test_microfacet_pdf_normalised(kind, rough):
    """The sampling pdf integrates to at most one over the hemisphere and the
    sampled directions follow it (E[1/pdf] over samples ~ visible solid angle)."""
    alb = np.array([0.8, 0.8, 0.8], np.float32)
    n = np.array([0, 0, 1], np.float32)
    wo = np.array([0.3, 0.0, np.sqrt(1 - 0.09)], np.float32)
    rng = np.random.default_rng(4)
    out = np.zeros(4, np.float32)
    # uniform hemisphere estimate of the integral of the pdf
    m = 20000
    z = rng.random(m)
    phi = 2 * np.pi * rng.random(m)
    r = np.sqrt(1 - z * z)
    acc = 0.0
    for k in range(m):
        wi = np.array([r[k] * np.cos(phi[k]), r[k] * np.sin(phi[k]), z[k]], np.float32)
        L.orc_brdf_eval(kind, ptr(alb), rough, ptr(wo), ptr(wi), ptr(n), ptr(out))
        acc += out[3]
    integral = acc / m * 2 * np.pi
    # the same integral is the probability that a sampled direction stays above the horizon
    smp = np.zeros(7, np.float32)
    valid = 0
    k = 20000
    for u in rng.random((k, 2)).astype(np.float32):
        L.orc_brdf_sample(kind, ptr(alb), rough, ptr(wo), ptr(n), float(u[0]), float(u[1]), ptr(smp))
        valid += smp[3] > 0
    assert integral <= 1.02
    assert abs(integral - valid / k) < 0.03


# ---- geometry -----------------------------------------------------------------------------
def test_cornell_scene_layout():
    tris = np.zeros((64, 16), np.float32)
    n = L.orc_cornell_triangles(ptr(tris))
    assert n == 32
    tris = tris[:n]
    em = tris[:, 12:15]
    assert (em.max(axis=1) > 0).sum() == 2 and np.all(em[em.max(axis=1) > 0] == 70.0)
    # light quad area 0.130 x 0.105
    area = sum(0.5 * np.linalg.norm(np.cross(t[3:6] - t[0:3], t[6:9] - t[0:3])) for t in tris[:2])
    assert abs(area - 0.13 * 0.105) < 1e-6
    # room walls face the room centre; box faces face away from the box centre
    center = np.array([0.278, 0.2744, 0.2796])
    for t in tris[:12]:
        c = (t[0:3] + t[3:6] + t[6:9]) / 3
        assert np.dot(t[9:12], center - c) > 0
    for t in tris:
        assert abs(np.linalg.norm(t[9:12]) - 1) < 1e-5


def test_bvh_closest_hit_equals_brute_force():
    rng = np.random.default_rng(9)
    m = 5000
    o = np.column_stack([rng.uniform(0.0, 0.556, m), rng.uniform(0.0, 0.548, m), rng.uniform(-0.8, 0.55, m)])
    d = rng.normal(size=(m, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays = np.ascontiguousarray(np.column_stack([o, d]), np.float32)
    a = np.zeros(3 * m, np.float32)
    b = np.zeros(3 * m, np.float32)
    L.orc_cornell_hits(ptr(rays), m, ptr(a), ptr(b))
    a, b = a.reshape(m, 3), b.reshape(m, 3)
    np.testing.assert_array_equal(a[:, 0], b[:, 0])
    hit = a[:, 0] > 0
    np.testing.assert_array_equal(a[hit, 1], b[hit, 1])
    assert hit.mean() > 0.3


# ---- pipeline invariants ------------------------------------------------------------------
def _render(**kw):
    rc = ob.config_c0(24, 20)
    for k, v in kw.items():
        setattr(rc, k, v)
    j = ob.OracleJob(rc, 1234, nthreads=2, skip_dead=True)
    j.render()
    return j


def test_rpd_off_keeps_raw_bytes_identical():
    """paper_eval.sh:195-196: RPD scale only changes reconstruction."""
    on, off = _render(rpf_shrinkage_scale=1.0), _render(rpf_shrinkage_scale=0.0)
    assert on.get("raw_radiance").tobytes() == off.get("raw_radiance").tobytes()
    assert np.array_equal(on.get("sample_count"), off.get("sample_count"))
    assert np.all(off.get("sensitivity_gradient") == 0) and np.all(off.get("sensitivity_confidence") == 0)


def test_dead_calls_do_not_change_outputs():
    rc = ob.config_c0(20, 16)
    a = ob.OracleJob(rc, 99, nthreads=1, skip_dead=False)
    b = ob.OracleJob(rc, 99, nthreads=1, skip_dead=True)
    a.render()
    b.render()
    for name in ("radiance", "raw_radiance", "uncertainty", "alpha", "variance_denoised", "adaptive_importance",
                 "sensitivity_gradient", "sample_count"):
        assert np.array_equal(a.get(name), b.get(name)), name


def test_budget_and_floor_of_adaptive_counts():
    rc = ob.config_c0(24, 20)
    rc.adaptive_base_samples = 1
    j = ob.OracleJob(rc, 7, nthreads=1)
    j.run_pass(np.full(24 * 20, rc.spp, np.int32))
    j.variance()
    c = j.adaptive_counts()
    assert c.min() >= 1
    assert c.sum() <= rc.adaptive_spp * 24 * 20
    assert c.sum() >= rc.adaptive_spp * 24 * 20 - 1
    assert c.max() <= max(4, 8 * rc.adaptive_spp) + 1


def test_no_hit_pixels_keep_mean_and_alpha_one():
    rc = ob.config_c0(24, 20)
    rc.cam_dir = (0.0, 0.0, -1.0)      # looks out of the open box: every primary ray misses
    rc.adaptive_passes = 0
    j = ob.OracleJob(rc, 3, nthreads=1)
    j.render()
    assert np.all(j.get("hit_count") == 0)
    assert np.all(j.get("alpha") == 1.0)
    assert np.all(j.get("radiance") == 0.0)


def test_record_fold_equals_traced_fold():
    """The oracle's fold of given records (used to check denoising of records)
    leaves exactly the state its own traced StatMC pass leaves."""
    rc = ob.config_c0(24, 20)
    rc.aperture = 0.14
    rc.focus_distance = 1.05
    counts = np.full(24 * 20, 3, np.int32)
    a = ob.OracleJob(rc, 77, nthreads=1, keep_records=True)
    a.run_pass(counts, statmc=True)
    b = ob.OracleJob(rc, 77)
    b.accumulate_records(counts, a.get("records"))
    for name in ("stats", "grid", "hit_count", "sample_count", "radiance", "normal", "albedo", "depth"):
        x, y = a.get(name), b.get(name)
        assert np.array_equal(x.view(np.uint8), y.view(np.uint8)), name
    a.recompute_sensitivity(); a.statmc()
    b.recompute_sensitivity(); b.statmc()
    assert np.array_equal(a.get("denoised"), b.get("denoised"))


def test_luminance_of_finite_channels_is_finite():
    """The grid gather adds its FP64 terms with fma(m, term, sum), m in {1, -0},
    which needs every term finite. Terms come from finite records, weights in
    [0, 1] and the float luminance 0.2126f r + 0.7152f g + 0.0722f b evaluated
    unfused left to right. Its float weights sum to exactly 1 and each rounded
    step is monotone in each channel, so the extremes are the all-FLT_MAX
    channels; those evaluate to FLT_MAX, not infinity."""
    from fractions import Fraction
    c = np.array([0.2126, 0.7152, 0.0722], np.float32)
    assert sum(Fraction(float(x)) for x in c) == 1
    m = np.finfo(np.float32).max
    for s in (1.0, -1.0):
        v = np.full(3, s * m, np.float32)
        with np.errstate(over="raise"):
            lum = (c[0] * v[0] + c[1] * v[1]) + c[2] * v[2]
        assert lum == np.float32(s) * m
    rng = np.random.default_rng(2)
    bits = (m.view(np.uint32) - rng.integers(0, 1 << 23, (200000, 3))).astype(np.uint32)
    v = bits.view(np.float32) * rng.choice(np.array([-1.0, 1.0], np.float32), (200000, 3))
    with np.errstate(over="raise"):
        lum = (c[0] * v[:, 0] + c[1] * v[:, 1]) + c[2] * v[:, 2]
    assert np.isfinite(lum).all()
