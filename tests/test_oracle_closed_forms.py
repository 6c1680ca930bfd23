"""Closed-form pins of the oracle's integrator, camera, StatMC filter and
sensitivity recompute (CPU only).

Each test builds a case whose answer follows from the mathematics alone, not
from the code under test:

* integrator (path_tracer/main.cpp:394-585): a Lambertian plane under a
  constant environment reflects rho * E; a mirror plane reflects exactly
  rho * E on every sample; in a white furnace (albedo 1 everywhere, no lights,
  a constant environment, an open enclosure) every path's expectation is E
  whatever the number of bounces, so Russian roulette's 1/p_rr and the MIS
  weights of environment and BRDF sampling are pinned; a rectangular light over
  a Lambertian floor gives rho * Le * F(p), F the form factor of a parallel
  rectangle (checked against numerical quadrature), which pins light sampling,
  its solid-angle pdf and the MIS weights of light and BRDF sampling;
* camera (core/camera.h:31-88, main.cpp:612-627): a pinhole maps the film
  linearly in tan-space with the vertical fov, the centre ray is the forward
  direction and the canonical Cornell image has the red wall on the left; a
  thin lens sends every ray of a film point through one point of the focus
  plane, from lens points spread uniformly over a disk of radius aperture / 2;
* StatMC (main.cpp:957-1127): on a constant image every neighbour has t = 0,
  so the weights are 1 / (1 + d^2) and alpha, the blend and the uncertainty
  are sums over the window; neighbours far apart in colour, or past the
  0.12 albedo gate, drop out of those sums;
* sensitivity (main.cpp:741-953, core/stats.h:93-303): radiance independent
  of every random dimension gives components near 0; radiance that is a
  function of one dimension's bin makes that component the dominant one; the
  node arithmetic is translation-equivariant.

Monte Carlo checks use the sample standard error of the estimate: |mean -
expected| <= 5 SE (a dropped term or a wrong weight moves the mean by many SE).
"""
import ctypes as C
import math

import numpy as np
import pytest
from scipy import integrate
from scipy import stats as sps

import oracle_bind as ob
import rpd_engine as eng

Z_LIMIT = 5.0


# ---- scene helpers ----------------------------------------------------------
def _scene(tris, mats, lights=(), env=(0.0, 0.0, 0.0)):
    T = (eng.Triangle * max(1, len(tris)))()
    for k, t in enumerate(tris):
        T[k] = t
    M = (eng.Material * len(mats))()
    for k, m in enumerate(mats):
        M[k] = m
    Lg = (eng.Light * max(1, len(lights)))()
    for k, l in enumerate(lights):
        Lg[k] = l
    S = (eng.Sphere * 1)()
    E = (C.c_float * 3)(*env)
    return T, S, M, Lg, E, [len(tris), 0, len(mats), len(lights)]


def _tri(v0, v1, v2, normal, material=0, emission=(0.0, 0.0, 0.0), light=-1):
    t = eng.Triangle()
    for name, v in (("v0", v0), ("v1", v1), ("v2", v2), ("normal", normal), ("emission", emission)):
        arr = getattr(t, name)
        for i in range(3):
            arr[i] = v[i]
    t.material_id = material
    t.light_id = light
    return t


def _quad_y(x0, x1, z0, z1, y, normal, material=0, emission=(0.0, 0.0, 0.0), light=-1):
    a, b, c, d = (x0, y, z0), (x1, y, z0), (x1, y, z1), (x0, y, z1)
    return [_tri(a, b, c, normal, material, emission, light), _tri(a, c, d, normal, material, emission, light)]


def _material(kind, albedo):
    m = eng.Material()
    m.kind = kind
    for i in range(3):
        m.albedo[i] = albedo[i]
        m.n1[i] = 1.0
        m.n2[i] = 1.5
    m.roughness = 0.5
    return m


def _rect_light(corner, width, length, emission):
    l = eng.Light()
    l.kind = eng.RPD_LIGHT_RECT
    for name, v in (("p0", corner), ("p1", width), ("p2", length), ("emission", emission)):
        arr = getattr(l, name)
        for i in range(3):
            arr[i] = v[i]
    return l


def _trace(rc, scene, seed=99):
    """One StatMC-capturing pass of the oracle; its per-sample records."""
    oj = ob.OracleJob(rc, seed, nthreads=4, keep_records=True, scene=scene)
    oj.run_pass(np.full(rc.width * rc.height, rc.spp, np.int32), statmc=True)
    return oj.get("records")


def _z(values, expected):
    """Standard score of the mean of `values` against `expected`."""
    v = np.asarray(values, np.float64)
    se = v.std(ddof=1) / math.sqrt(len(v))
    return (v.mean() - expected) / max(se, 1e-30), v.mean(), se


# ---- integrator ---------------------------------------------------------------
def _looking_down(w=48, h=48, spp=8):
    return ob.RenderConfig(width=w, height=h, spp=spp, fov=40.0, cam_pos=(0.0, 2.0, 0.0), cam_dir=(0.0, -1.0, 0.0),
                           cam_up=(0.0, 0.0, 1.0), focus_distance=2.0)


def test_lambert_plane_under_constant_environment_reflects_rho_e():
    rho, env = (0.25, 0.5, 0.8), (0.6, 1.0, 2.0)
    scene = _scene(_quad_y(-50, 50, -50, 50, 0.0, (0, 1, 0)), [_material(eng.RPD_MAT_LAMBERTIAN, rho)], env=env)
    rec = _trace(_looking_down(), scene)
    assert (rec[:, 5] == 1).all()                               # every camera ray hits the plane
    for c in range(3):
        z, mean, se = _z(rec[:, 1 + c], rho[c] * env[c])
        print(f"channel {c}: mean {mean:.5f} expected {rho[c] * env[c]:.5f} (z {z:+.2f})")
        assert abs(z) <= Z_LIMIT


def test_mirror_plane_reflects_rho_e_on_every_sample():
    rho, env = (0.3, 0.6, 0.9), (1.5, 1.0, 0.5)
    scene = _scene(_quad_y(-50, 50, -50, 50, 0.0, (0, 1, 0)), [_material(eng.RPD_MAT_MIRROR, rho)], env=env)
    rec = _trace(_looking_down(16, 16, 4), scene)
    np.testing.assert_allclose(rec[:, 1:4], np.tile(np.float32(rho) * np.float32(env), (len(rec), 1)), rtol=1e-6)


def _furnace_cornell(env):
    """cornell_box's geometry with albedo 1 on every surface and no light."""
    tris, _, mats, _, _, n = eng.builtin_scene_arrays("cornell_box")
    out = []
    for k in range(n[0]):
        t = tris[k]
        for i in range(3):
            t.emission[i] = 0.0
        t.light_id = -1
        out.append(t)
    white = [_material(eng.RPD_MAT_LAMBERTIAN, (1.0, 1.0, 1.0)) for _ in range(n[2])]
    return _scene(out, white, env=env)


def test_white_furnace_every_path_expects_the_environment():
    """Albedo 1 makes each bounce's weight 1, so a path that escapes through
    the open front of the box after k Russian-roulette survivals contributes
    E / 0.8^k with probability 0.8^k: the expectation is E for every pixel."""
    env = (0.5, 1.0, 1.5)
    rc = ob.config_c0(48, 48)
    rc.spp = 16
    rec = _trace(rc, _furnace_cornell(env))
    bounces = (rec[:, 28].astype(np.int64) & 64) != 0
    assert bounces.mean() > 0.5                                   # most paths reach a Russian-roulette test
    for c in range(3):
        z, mean, se = _z(rec[:, 1 + c], env[c])
        print(f"channel {c}: mean {mean:.5f} expected {env[c]} se {se:.5f} (z {z:+.2f})")
        assert abs(z) <= Z_LIMIT
    # also within each image quarter (deep corners and the open front alike)
    px = rec[:, 0].astype(np.int64)
    quarter = (px % rc.width >= rc.width // 2).astype(int) + 2 * (px // rc.width >= rc.height // 2)
    for q in range(4):
        z, _, _ = _z(rec[quarter == q, 2], env[1])
        assert abs(z) <= Z_LIMIT, (q, z)


def _parallel_rect_form_factor(px, pz, h, x0, x1, z0, z1):
    """Form factor from a floor element at (px, 0, pz), facing +y, to the
    rectangle [x0,x1] x [z0,z1] at height h facing -y: inclusion-exclusion of
    corner rectangles, G the standard form factor of a parallel rectangle with
    a corner above the element (Howell's catalogue, configuration B-3)."""
    def G(a, b):
        X, Y = a / h, b / h
        sx, sy = np.sqrt(1.0 + X * X), np.sqrt(1.0 + Y * Y)
        return (X / sx * np.arctan(Y / sx) + Y / sy * np.arctan(X / sy)) / (2.0 * np.pi)
    return G(x1 - px, z1 - pz) - G(x0 - px, z1 - pz) - G(x1 - px, z0 - pz) + G(x0 - px, z0 - pz)


def test_parallel_rect_form_factor_matches_quadrature():
    h, rect = 0.7, (-0.3, 0.5, -0.2, 0.4)
    for px, pz in [(0.0, 0.0), (0.9, -0.6), (-0.25, 0.35)]:
        f = lambda z, x: h * h / (np.pi * ((x - px) ** 2 + (z - pz) ** 2 + h * h) ** 2)  # cos cos / (pi r^2)
        q, _ = integrate.dblquad(f, rect[0], rect[1], rect[2], rect[3], epsabs=1e-11)
        assert abs(_parallel_rect_form_factor(px, pz, h, *rect) - q) < 1e-8


def test_rect_light_over_lambert_floor_gives_rho_le_times_form_factor():
    rho, le, h = (0.8, 0.5, 0.3), (4.0, 6.0, 8.0), 0.6
    x0, x1, z0, z1 = -0.4, 0.3, -0.2, 0.5
    tris = _quad_y(-3, 3, -3, 3, 0.0, (0, 1, 0), material=0)
    tris += _quad_y(x0, x1, z0, z1, h, (0, -1, 0), material=1, emission=le, light=0)
    # width x length points down (-y), toward the floor
    light = _rect_light((x0, h, z0), (x1 - x0, 0, 0), (0, 0, z1 - z0), le)
    scene = _scene(tris, [_material(eng.RPD_MAT_LAMBERTIAN, rho), _material(eng.RPD_MAT_LAMBERTIAN, (0.5,) * 3)],
                   lights=[light])
    rc = ob.RenderConfig(width=48, height=32, spp=16, fov=50.0, cam_pos=(0.0, 0.3, -1.5),
                         cam_dir=(0.0, -0.1, 1.0), focus_distance=2.0)
    rec = _trace(rc, scene)
    floor = (rec[:, 5] == 1) & (np.abs(rec[:, 13]) < 1e-4)
    on_light = (rec[:, 5] == 1) & (np.abs(rec[:, 13] - h) < 1e-4)
    assert floor.sum() > 5000 and on_light.sum() > 50
    # the light's emitting side, seen from below: Le exactly
    np.testing.assert_array_equal(rec[on_light, 1:4], np.tile(np.float32(le), (on_light.sum(), 1)))
    F = _parallel_rect_form_factor(rec[floor, 12].astype(np.float64), rec[floor, 14].astype(np.float64), h,
                                   x0, x1, z0, z1)
    for c in range(3):
        resid = rec[floor, 1 + c] - rho[c] * le[c] * F
        z, mean, se = _z(resid, 0.0)
        print(f"channel {c}: mean residual {mean:+.2e} se {se:.2e} (z {z:+.2f}), mean expected "
              f"{float((rho[c] * le[c] * F).mean()):.4f}")
        assert abs(z) <= Z_LIMIT


# ---- camera -------------------------------------------------------------------
def _film_coords(rec, w, h):
    px = rec[:, 0].astype(np.int64)
    u = ((px % w) + rec[:, 16].astype(np.float64)) / w
    v = 1.0 - ((px // w) + rec[:, 17].astype(np.float64)) / h
    return u, v


def _basis(rc):
    fwd = np.asarray(rc.cam_dir, np.float64)
    fwd /= np.linalg.norm(fwd)
    right = np.cross(fwd, np.asarray(rc.cam_up, np.float64))
    right /= np.linalg.norm(right)
    up = np.cross(right, fwd)
    return fwd, right, up


def test_pinhole_projection_is_linear_in_tan_space_with_the_vertical_fov():
    rc = ob.config_c0(40, 30)
    rec = _trace(rc, None)                    # the built-in cornell_box
    hit = rec[:, 5] == 1
    rec = rec[hit]
    d = rec[:, 12:15].astype(np.float64) - np.asarray(rc.cam_pos)
    fwd, right, up = _basis(rc)
    u, v = _film_coords(rec, rc.width, rc.height)
    t = math.tan(math.radians(rc.fov) / 2.0)
    aspect = rc.width / rc.height
    along = d @ fwd
    np.testing.assert_allclose((d @ right) / along, (2.0 * u - 1.0) * aspect * t, atol=2e-5)
    np.testing.assert_allclose((d @ up) / along, (2.0 * v - 1.0) * t, atol=2e-5)
    # the canonical Cornell image: red wall (x = 0.5528 side) on the left, green on the right
    red = (rec[:, 9] > 0.5) & (rec[:, 10] < 0.2)
    green = (rec[:, 10] > 0.4) & (rec[:, 9] < 0.2)
    lu = u[red].mean()
    ru = u[green].mean()
    assert lu < 0.25 and ru > 0.75, (lu, ru)


def test_thin_lens_rays_of_one_film_point_meet_on_the_focus_plane():
    rc = ob.config_c1(40, 30)                  # aperture 0.14, focus distance 1.05
    rec = _trace(rc, None)
    rec = rec[rec[:, 5] == 1]
    fwd, right, up = _basis(rc)
    cam = np.asarray(rc.cam_pos, np.float64)
    u, v = _film_coords(rec, rc.width, rc.height)
    t = math.tan(math.radians(rc.fov) / 2.0)
    aspect = rc.width / rc.height
    focus = cam + rc.focus_distance * (fwd + ((2 * u - 1) * aspect * t)[:, None] * right
                                       + ((2 * v - 1) * t)[:, None] * up)
    hitp = rec[:, 12:15].astype(np.float64)
    # the ray through the focus point and the hit crosses the lens plane (through
    # the camera, normal to fwd) at the lens point it started from
    dirn = hitp - focus
    s = ((cam - focus) @ fwd) / (dirn @ fwd)
    lens = focus + s[:, None] * dirn - cam
    R = rc.aperture / 2.0
    r2 = ((lens @ right) ** 2 + (lens @ up) ** 2) / (R * R)
    assert r2.max() <= 1.0 + 1e-3
    # lens points are uniform on the disk: r^2 / R^2 ~ U(0, 1), angle ~ U(0, 2 pi)
    assert sps.kstest(np.clip(r2, 0, 1), "uniform").pvalue > 1e-3
    ang = (np.arctan2(lens @ up, lens @ right) + 2 * np.pi) % (2 * np.pi)
    assert sps.kstest(ang / (2 * np.pi), "uniform").pvalue > 1e-3


# ---- StatMC ---------------------------------------------------------------------
def _statmc_job(w, h, radius=4, **cfg):
    rc = ob.config_c0(w, h)
    rc.color_window_radius = radius
    rc.rpf_shrinkage_scale = 0.0               # no sensitivity relaxation: q is the Welch statistic alone
    for k, v in cfg.items():
        setattr(rc, k, v)
    return ob.OracleJob(rc, 5, nthreads=1), rc


def _set_pixels(oj, w, h, mean, var, n, albedo, normal=(0.0, 0.0, 1.0), depth=3.0):
    """Per-pixel statistics of n samples with luminance mean `mean` (H, W) and
    sample variance `var`, colour mean grey; sqrt-domain statistics from them;
    features from uniform sums over n hits."""
    N = w * h
    mean = np.broadcast_to(np.asarray(mean, np.float32), (h, w)).ravel()
    stats = np.zeros((N, 12), np.float32)
    stats[:, 0] = mean
    stats[:, 1] = var * (n - 1)
    stats[:, 2] = np.full(N, n, np.int32).view(np.float32)
    stats[:, 3:6] = mean[:, None]
    stats[:, 6:9] = np.sqrt(mean)[:, None]
    stats[:, 9:12] = var / (4.0 * mean[:, None]) * (n - 1)    # delta-method variance of sqrt(c)
    oj.set("stats", stats)
    oj.set("hit_count", np.full(N, n, np.int32))
    oj.set("sample_count", np.full(N, n, np.int32))
    oj.set("normal", np.tile(np.float32(normal) * n, (N, 1)))
    oj.set("depth", np.full(N, depth * n, np.float32))
    alb = np.broadcast_to(np.asarray(albedo, np.float32), (h, w, 3)).reshape(N, 3)
    oj.set("albedo", alb * n)
    oj.recompute_sensitivity()                 # zero sensitivity (rpf_shrinkage_scale 0)


def _window_sums(mask_fn, w, h, r):
    """Sum of 1/(1+d^2) and of its square over accepted neighbours (mask_fn(x,
    y, xx, yy) true), with the count of accepted non-centre neighbours."""
    S1 = np.zeros((h, w))
    S2 = np.zeros((h, w))
    A = np.zeros((h, w))
    for y in range(h):
        for x in range(w):
            for dy in range(-r, r + 1):
                for dx in range(-r, r + 1):
                    xx, yy = x + dx, y + dy
                    if not (0 <= xx < w and 0 <= yy < h) or not mask_fn(x, y, xx, yy):
                        continue
                    wt = 1.0 / (1.0 + dx * dx + dy * dy)
                    S1[y, x] += wt
                    S2[y, x] += wt * wt
                    if dx or dy:
                        A[y, x] += 1
    return S1, S2, A


def _expected_alpha_unc(S1, S2, A, vm):
    neff = S1 * S1 / np.maximum(1e-8, S2)
    blend = np.where((A >= 2) & (S1 >= 1.1), np.clip(1.0 - 1.0 / np.sqrt(neff), 0.0, 0.8), 0.0)
    a = 1.0 - blend
    unc = a * a * vm + blend * blend * vm * S2 / (S1 * S1) + 2.0 * a * blend * vm / S1
    return a, unc


def test_statmc_constant_image_closed_form():
    w, h, r, n, var = 23, 17, 4, 16, 0.09
    oj, rc = _statmc_job(w, h, r)
    _set_pixels(oj, w, h, 0.5, var, n, (0.4, 0.4, 0.4))
    oj.statmc()
    S1, S2, A = _window_sums(lambda *a: True, w, h, r)
    a, unc = _expected_alpha_unc(S1, S2, A, var / n)
    np.testing.assert_allclose(oj.get("alpha").reshape(h, w), a, rtol=1e-6)
    np.testing.assert_allclose(oj.get("uncertainty").reshape(h, w), unc, rtol=1e-5)
    np.testing.assert_allclose(oj.get("denoised"), 0.5, rtol=1e-6)
    # both branches of the blend: clamped at 0.8 inside, 1 - 1/sqrt(n_eff) in the corners
    blend = 1.0 - a
    assert blend[h // 2, w // 2] == pytest.approx(0.8) and 0.3 < blend[0, 0] < 0.79


def test_statmc_rejects_an_incompatible_half():
    """Left half mean 0.2, right half 5.0 with small variance: across the
    boundary t^2 is in the thousands, so q > 25 rejects every cross pair and
    each pixel filters over its own half of the window only."""
    w, h, r, n, var = 20, 12, 3, 32, 0.01
    oj, rc = _statmc_job(w, h, r)
    mean = np.where(np.arange(w)[None, :] < w // 2, 0.2, 5.0) * np.ones((h, 1))
    _set_pixels(oj, w, h, mean, var, n, (0.4, 0.4, 0.4))
    oj.statmc()
    side = lambda x, y, xx, yy: (x < w // 2) == (xx < w // 2)
    S1, S2, A = _window_sums(side, w, h, r)
    a, unc = _expected_alpha_unc(S1, S2, A, var / n)
    np.testing.assert_allclose(oj.get("alpha").reshape(h, w), a, rtol=1e-6)
    np.testing.assert_allclose(oj.get("uncertainty").reshape(h, w), unc, rtol=1e-5)
    np.testing.assert_allclose(oj.get("denoised")[:, 0].reshape(h, w), mean, rtol=1e-6)


@pytest.mark.parametrize("delta,rejected", [(0.11 / math.sqrt(3), False), (0.13 / math.sqrt(3), True)])
def test_statmc_albedo_gate_at_0_12(delta, rejected):
    """A checkerboard of albedo grey levels a and a + delta: the L2 distance of
    neighbours of opposite colour is delta * sqrt(3), accepted at 0.11 and
    rejected at 0.13 (the 0.12 gate, core/stats.h:45)."""
    w, h, r, n, var = 15, 13, 3, 16, 0.05
    oj, rc = _statmc_job(w, h, r)
    par = (np.arange(w)[None, :] + np.arange(h)[:, None]) % 2
    alb = (0.3 + delta * par)[:, :, None] * np.ones(3)
    _set_pixels(oj, w, h, 0.7, var, n, alb)
    oj.statmc()
    mask = (lambda x, y, xx, yy: (x + y) % 2 == (xx + yy) % 2) if rejected else (lambda *a: True)
    S1, S2, A = _window_sums(mask, w, h, r)
    a, _ = _expected_alpha_unc(S1, S2, A, var / n)
    np.testing.assert_allclose(oj.get("alpha").reshape(h, w), a, rtol=1e-6)


# ---- sensitivity ------------------------------------------------------------------
DIMS = {  # component tile field, record columns / flag bit written by the generator
    "brdf": 4, "lens": 6, "light_uv": 10, "light_select": 12, "environment": 14, "rr": 16}
RELIABLE = ("brdf", "lens", "light_uv", "light_select", "environment", "rr")


def _records(w, h, spp, rng, depends_on=None, region=None):
    """Records of a plain image: every dimension valid and uniform, radiance
    lognormal; with depends_on, luminance rises with that dimension's bin (in
    `region` only: a predicate on (x, y), all pixels when None)."""
    N = w * h
    n = N * spp
    rec = np.zeros((n, 32), np.float32)
    pix = np.repeat(np.arange(N), spp)
    rec[:, 0] = pix
    rec[:, 4] = 1.0
    rec[:, 5] = 1.0
    rec[:, 6:9] = (0.0, 0.0, 1.0)
    rec[:, 9:12] = 0.5
    rec[:, 12:15] = 0.0
    rec[:, 15] = 4.0
    rec[:, 16:26] = rng.random((n, 10), dtype=np.float32)
    rec[:, 26] = rng.integers(0, 16, n)
    rec[:, 27] = rng.random(n) < 0.5
    rec[:, 28] = 1 | 2 | 4 | 8 | 16 | 32 | 64
    y = rng.lognormal(-1.0, 0.3, n)
    if depends_on is not None:
        if depends_on == "light_select":
            level = rec[:, 26] % 4
        elif depends_on == "rr":
            level = 3.0 * rec[:, 27]
        else:
            col = {"brdf": 20, "lens": 18, "light_uv": 22, "environment": 24}[depends_on]
            level = np.clip((rec[:, col] * 4).astype(int), 0, 3)
        on = np.ones(n, bool) if region is None else region(pix % w, pix // w)
        y = np.where(on, y + 2.0 * level, y)
    rec[:, 1:4] = y[:, None]
    return rec


def _sensitivity_tiles(w, h, spp, rec, n0=32.0):
    rc = ob.config_c0(w, h)
    rc.rpf_confidence_samples = n0
    oj = ob.OracleJob(rc, 5, nthreads=1)
    oj.accumulate_records(np.full(w * h, spp, np.int32), rec)
    oj.recompute_sensitivity()
    t = oj.get("tiles")
    idx = {"all": 0, "pixel": 2, "brdf": 4, "lens": 6, "light": 8, "light_uv": 10, "light_select": 12,
           "environment": 14, "rr": 16}
    return {k: t[v].reshape(oj.nodes_y, oj.nodes_x) for k, v in idx.items()}, oj


def test_sensitivity_of_radiance_varying_only_with_screen_position_is_near_zero():
    """Luminance that changes with the pixel position and with no random
    dimension: every random component is small against its screen-position
    counterpart, so the relative (random / (random + screen)) and reliable
    components are near 0, while the screen-position dependency is high."""
    w = h = 32
    rng = np.random.default_rng(1)
    rec = _records(w, h, 32, rng)
    pix = rec[:, 0].astype(np.int64)
    rec[:, 1:4] = (0.2 + 0.5 * ((pix % w) % 4) + 0.3 * ((pix // w) % 4))[:, None]
    rec[:, 1:4] *= rng.lognormal(0.0, 0.05, len(rec))[:, None].astype(np.float32)
    tiles, _ = _sensitivity_tiles(w, h, 32, rec)
    # the noise floor: sqrt of a bias-corrected explained variance of about
    # bins / samples (16 / ~2000 per node), against >= 0.5 for a real dependence
    for k in RELIABLE + ("all",):
        print(k, float(tiles[k].mean()), float(tiles[k].max()))
        assert tiles[k].mean() < 0.15 and tiles[k].max() < 0.25, (k, tiles[k].mean(), tiles[k].max())
    # the screen-position dependency itself (the pixel accumulator, not normalised) is high
    assert tiles["pixel"][1:-1, 1:-1].min() > 0.5


@pytest.mark.parametrize("dim", RELIABLE)
def test_sensitivity_dependence_on_one_dimension_dominates(dim):
    """Luminance = a screen-position term + a term rising with one dimension's
    bin. Each random component is normalised by the screen dependency of the
    same samples (random / (random + screen)), so the dependent one is about
    r / (r + s) ~ 0.5 and the others sit at the noise floor: the dependent
    component is the dominant one (the `all` field) at every node."""
    w = h = 32
    rng = np.random.default_rng(2)
    rec = _records(w, h, 32, rng, depends_on=dim)
    pix = rec[:, 0].astype(np.int64)
    rec[:, 1:4] += (1.0 * ((pix % w) % 4) + 0.6 * ((pix // w) % 4))[:, None].astype(np.float32)
    tiles, _ = _sensitivity_tiles(w, h, 32, rec)
    inner = (slice(1, -1), slice(1, -1))
    print(dim, {k: round(float(tiles[k][inner].mean()), 3) for k in RELIABLE})
    # measured: the dependent component 0.62-0.70, the others at most 0.27
    assert tiles[dim][inner].min() > 0.5, (dim, tiles[dim][inner].min())
    for k in RELIABLE:
        if k != dim:
            assert tiles[k][inner].max() < 0.35, (dim, k, tiles[k][inner].max())
    np.testing.assert_array_equal(tiles["all"][inner], tiles[dim][inner])      # the dominant component
    if dim in ("light_uv", "light_select"):
        np.testing.assert_array_equal(tiles["light"][inner], tiles[dim][inner])


def test_sensitivity_node_arithmetic_is_translation_equivariant():
    """The same pixels' records shifted right by one grid step give the same
    node values shifted by one node, away from the borders."""
    w, h, spp = 40, 24, 16
    rc = ob.config_c0(w, h)
    step = ob.OracleJob(rc, 5, nthreads=1).step
    rng = np.random.default_rng(3)
    region = lambda x, y: (x >= 12) & (x < 24)
    base = _records(w, h, spp, rng, depends_on="brdf", region=region)
    # shifted image: pixel (x, y) takes the records of pixel (x - step, y)
    x, y = np.arange(w * h) % w, np.arange(w * h) // w
    shifted = base.reshape(w * h, spp, 32)[y * w + (x - step) % w].reshape(-1, 32).copy()
    shifted[:, 0] = np.repeat(np.arange(w * h), spp)
    t0, oj = _sensitivity_tiles(w, h, spp, base)
    t1, _ = _sensitivity_tiles(w, h, spp, shifted)
    for k in ("brdf", "all"):
        a, b = t0[k], t1[k]
        assert np.ptp(a[2:-2, 2:-4]) > 0.3                # a varied field, not a constant one
        np.testing.assert_allclose(a[2:-2, 2:-4], b[2:-2, 3:-3], rtol=1e-6, atol=1e-7, err_msg=k)


def _block_records(w, h, spp, rng, block_x):
    """Plain records plus a screen term; pixels with x in block_x also depend
    on brdf_u's bin."""
    rec = _records(w, h, spp, rng, depends_on="brdf", region=lambda x, y: (x >= block_x[0]) & (x < block_x[1]))
    pix = rec[:, 0].astype(np.int64)
    rec[:, 1:4] += (1.0 * ((pix % w) % 4) + 0.6 * ((pix // w) % 4))[:, None].astype(np.float32)
    return rec


def test_sensitivity_prior_blends_toward_the_raw_value_with_confidence():
    """final = lerp(prior, raw, confidence) (main.cpp:868-894). Records that
    change only pixels splatted into the right neighbour of node a (x in
    [(a+1) step, (a+2) step)) leave node a's raw value alone: at confidence
    ~1 (rpf_confidence_samples tiny) node a keeps its value, at confidence
    ~0 (huge) it follows its 3x3 prior and moves."""
    w, h, spp = 48, 24, 16
    step = ob.OracleJob(ob.config_c0(w, h), 5, nthreads=1).step
    a = 5
    base = _block_records(w, h, spp, np.random.default_rng(4), (0, 0))
    changed = _block_records(w, h, spp, np.random.default_rng(4), ((a + 1) * step, (a + 2) * step))
    for n0, moves in [(1e-6, False), (1e9, True)]:
        t0, _ = _sensitivity_tiles(w, h, spp, base, n0=n0)
        t1, _ = _sensitivity_tiles(w, h, spp, changed, n0=n0)
        rows = slice(1, -1)
        # every record of pixels that splat into node a is unchanged
        np.testing.assert_array_equal(base[(base[:, 0] % w) < (a + 1) * step], changed[(changed[:, 0] % w) < (a + 1) * step])
        for k in ("brdf", "pixel"):
            d = np.abs(t1[k][rows, a] - t0[k][rows, a])
            d_next = np.abs(t1[k][rows, a + 1] - t0[k][rows, a + 1])
            print(f"n0 {n0:g} {k}: node a moved {d.max():.2e}, node a+1 moved {d_next.max():.2e}")
            if moves:
                assert d.max() > 0.3 * d_next.max() > 0
            else:
                assert d.max() < 1e-5 < d_next.max()


def test_sensitivity_node_arithmetic_is_transpose_equivariant():
    """A square image's records with x and y exchanged give the transposed
    tiles: bins are relabelled (explained variance does not depend on the
    labels), and the splat weights, the 3x3 prior's 1 / (1 + dx^2 + dy^2)
    weights and the gradient are symmetric in x and y."""
    w = h = 32
    spp = 16
    rng = np.random.default_rng(6)
    rec = _records(w, h, spp, rng, depends_on="lens", region=lambda x, y: (x >= 8) & (x < 20) & (y >= 4) & (y < 12))
    pix = rec[:, 0].astype(np.int64)
    rec[:, 1:4] += (0.7 * ((pix % w) % 4) + 0.7 * ((pix // w) % 4))[:, None].astype(np.float32)
    x, y = np.arange(w * h) % w, np.arange(w * h) // w
    tr = rec.reshape(w * h, spp, 32)[x * w + y].reshape(-1, 32).copy()      # pixel (x, y) <- pixel (y, x)
    tr[:, 0] = np.repeat(np.arange(w * h), spp)
    tr[:, [16, 17]] = tr[:, [17, 16]]                                        # the sub-pixel offset is transposed too
    t0, _ = _sensitivity_tiles(w, h, spp, rec)
    t1, _ = _sensitivity_tiles(w, h, spp, tr)
    for k in ("lens", "pixel", "all"):
        assert np.ptp(t0[k]) > 0.3
        np.testing.assert_allclose(t1[k], t0[k].T, rtol=1e-5, atol=1e-6, err_msg=k)


def test_sensitivity_normalises_each_component_by_its_own_screen_subset():
    """rpf::Sample validity: a component's screen comparator is the screen
    dependency of the samples where that dimension is valid (main.cpp:341-360).
    Here the lens is valid only on samples whose luminance ignores the screen
    position, the other samples vary with it: brdf (valid everywhere) sees a
    high screen dependency and stays at the noise floor; the lens's screen
    comparator is near 0 and its pixel (screen) field is high only through
    the lens-invalid samples."""
    w = h = 32
    spp = 32
    rng = np.random.default_rng(8)
    rec = _records(w, h, spp, rng)
    pix = rec[:, 0].astype(np.int64)
    lens_valid = rng.random(len(rec)) < 0.5
    rec[:, 28] = np.where(lens_valid, 127, 127 & ~2)
    screen = (1.0 * ((pix % w) % 4) + 0.6 * ((pix // w) % 4)).astype(np.float32)
    rec[~lens_valid, 1:4] += screen[~lens_valid, None]
    tiles, _ = _sensitivity_tiles(w, h, spp, rec)
    inner = (slice(1, -1), slice(1, -1))
    print("brdf", float(tiles["brdf"][inner].mean()), "lens", float(tiles["lens"][inner].mean()))
    assert tiles["brdf"][inner].max() < 0.35
    assert tiles["pixel"][inner].min() > 0.3
