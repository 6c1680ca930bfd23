"""Image-quality metrics (shared/tools/eval_tools.cpp:120-285, plus relMSE).

CPU: the oracle's restatement against
  * the reference's own self-test (eval_tools.cpp:317-338);
  * an independent numpy computation with direct 11x11 windows (no integral
    images) and explicit finite differences, on ragged image sizes where the
    border clipping matters;
  * closed forms on constant images.
GPU: the engine's device metrics against the oracle, interleaved host images
and a job's planar output buffer.
"""
import numpy as np
import pytest

import oracle_bind as ob


def _numpy_metrics(ref, tst):
    """The same metrics written independently in numpy (float64)."""
    a = ref.astype(np.float64)
    b = tst.astype(np.float64)
    h, w, _ = a.shape
    n = h * w
    d = a - b
    peak = np.abs(a).max()
    peak = 1.0 if peak <= 0 else peak
    out = {"mae": np.abs(d).mean(), "rmse": np.sqrt((d * d).mean()), "peak": peak}
    out["psnr"] = 20 * np.log10(peak / out["rmse"]) if out["rmse"] > 0 else np.inf
    C1, C2 = (0.01 * peak) ** 2, (0.03 * peak) ** 2
    ss = []
    for c in range(3):
        x, y = a[..., c].ravel(), b[..., c].ravel()
        mx, my = x.mean(), y.mean()
        vx, vy, cv = ((x - mx) ** 2).mean(), ((y - my) ** 2).mean(), ((x - mx) * (y - my)).mean()
        den = (mx * mx + my * my + C1) * (vx + vy + C2)
        ss.append((2 * mx * my + C1) * (2 * cv + C2) / den if den > 0 else 0.0)
    out.update(ssim_r=ss[0], ssim_g=ss[1], ssim_b=ss[2], ssim_mean=sum(ss) / 3)
    out["relmse"] = (d * d / (a * a + 1e-2)).mean()
    out["nrmse"] = np.sqrt((d * d).sum() / max(1e-18, (a * a).sum()))
    wts = np.array([0.2126, 0.7152, 0.0722])
    la = np.log1p(np.maximum(0.0, a @ wts))
    lb = np.log1p(np.maximum(0.0, b @ wts))
    lp = max(1e-6, la.max())
    c1, c2 = (0.01 * lp) ** 2, (0.03 * lp) ** 2
    tot = 0.0
    for yy in range(h):
        for xx in range(w):
            wa = la[max(0, yy - 5):min(h, yy + 6), max(0, xx - 5):min(w, xx + 6)]
            wb = lb[max(0, yy - 5):min(h, yy + 6), max(0, xx - 5):min(w, xx + 6)]
            ma, mb = wa.mean(), wb.mean()
            va = max(0.0, (wa * wa).mean() - ma * ma)
            vb = max(0.0, (wb * wb).mean() - mb * mb)
            cv = (wa * wb).mean() - ma * mb
            den = (ma * ma + mb * mb + c1) * (va + vb + c2)
            tot += (2 * ma * mb + c1) * (2 * cv + c2) / den if den > 0 else 1.0
    out["log_ssim11"] = tot / n
    gx_a, gy_a = la[:-1, 1:] - la[:-1, :-1], la[1:, :-1] - la[:-1, :-1]
    gx_b, gy_b = lb[:-1, 1:] - lb[:-1, :-1], lb[1:, :-1] - lb[:-1, :-1]
    ge = ((gx_a - gx_b) ** 2 + (gy_a - gy_b) ** 2).sum()
    gr = (gx_a ** 2 + gy_a ** 2).sum()
    out["log_gradient_nrmse"] = np.sqrt(ge / max(1e-18, gr))
    return out


def _close(got, want, rel=1e-10, abs_=1e-13):
    for k, v in want.items():
        g = got[k]
        if np.isinf(v):
            assert np.isinf(g), k
        else:
            assert abs(g - v) <= abs_ + rel * abs(v), (k, g, v)


def test_reference_self_test():
    """eval_tools.cpp:317-338: a 16x16 ramp against itself and with one value changed."""
    s = 16
    y, x = np.mgrid[0:s, 0:s]
    ref = np.repeat(((x + y + 1) / (2.0 * s)).astype(np.float32)[..., None], 3, axis=2)
    changed = ref.copy()
    changed[8, 8, 0] += np.float32(0.25)
    same = ob.image_metrics(ref, ref)
    diff = ob.image_metrics(ref, changed)
    assert same["nrmse"] <= 1e-12 and abs(same["log_ssim11"] - 1.0) <= 1e-12 and same["log_gradient_nrmse"] <= 1e-12
    assert diff["nrmse"] > 0 and diff["log_ssim11"] < 1.0 and diff["log_gradient_nrmse"] > 0


@pytest.mark.parametrize("h,w,seed", [(17, 23, 1), (9, 31, 2), (12, 5, 3), (1, 14, 4)])
def test_oracle_matches_direct_windows(h, w, seed):
    rng = np.random.default_rng(seed)
    ref = rng.lognormal(-1.0, 1.0, (h, w, 3)).astype(np.float32)
    tst = (ref * rng.lognormal(0.0, 0.3, (h, w, 3))).astype(np.float32)
    tst[0, 0] = -0.5   # negative luminance is clamped before log1p
    _close(ob.image_metrics(ref, tst), _numpy_metrics(ref, tst))


def test_closed_forms_on_constant_images():
    c, d = 0.75, 0.125
    ref = np.full((8, 10, 3), c, np.float32)
    tst = np.full((8, 10, 3), c + d, np.float32)
    m = ob.image_metrics(ref, tst)
    assert m["mae"] == pytest.approx(d, rel=1e-12)
    assert m["rmse"] == pytest.approx(d, rel=1e-12)
    assert m["psnr"] == pytest.approx(20 * np.log10(c / d), rel=1e-12)
    assert m["nrmse"] == pytest.approx(d / c, rel=1e-12)
    assert m["relmse"] == pytest.approx(d * d / (c * c + 1e-2), rel=1e-12)
    assert m["log_gradient_nrmse"] == 0.0
    C1 = (0.01 * c) ** 2
    want = (2 * c * (c + d) + C1) / (c * c + (c + d) ** 2 + C1)   # both variances are zero
    assert m["ssim_r"] == pytest.approx(want, rel=1e-12)


@pytest.mark.gpu
@pytest.mark.parametrize("h,w,seed", [(17, 23, 1), (64, 48, 2), (300, 257, 3)])
def test_device_metrics_match_oracle(engine_ctx, h, w, seed):
    import rpd_engine as eng
    rng = np.random.default_rng(seed)
    ref = rng.lognormal(-1.0, 1.0, (h, w, 3)).astype(np.float32)
    tst = (ref * rng.lognormal(0.0, 0.3, (h, w, 3))).astype(np.float32)
    want = ob.image_metrics(ref, tst)
    got = eng.image_metrics(engine_ctx, ref, tst)
    print({k: (got[k], want[k]) for k in want})
    _close(got, want, rel=1e-9, abs_=1e-12)
