"""The reference's comparison denoisers (denoising/joint_bilateral.cpp,
denoising/atrous_wavelet.cpp).

CPU: the oracle's float restatement against an independent float64 numpy
implementation, against limits (huge range sigmas turn the bilateral filters
into the Gaussian and the a-trous pass into the plain B3-spline smoothing,
which keeps a linear ramp away from the border) and on constant images.
GPU: the device filters against the oracle on the same inputs; only the float
exp/acos of the two math libraries differ, so 1e-5 relative.
"""
import numpy as np
import pytest

import oracle_bind as ob


def _inputs(h, w, seed):
    rng = np.random.default_rng(seed)
    raw = rng.lognormal(-1.0, 0.8, (h, w, 3)).astype(np.float32)
    nrm = rng.normal(0.0, 1.0, (h, w, 3))
    nrm[..., 2] = np.abs(nrm[..., 2]) + 1.0
    nrm /= np.linalg.norm(nrm, axis=2, keepdims=True)
    enc = (0.5 * nrm + 0.5).astype(np.float32)         # joint_bilateral decodes 2n - 1
    alb = rng.uniform(0.1, 0.9, (h, w, 3)).astype(np.float32)
    return raw, enc, alb


def _np_bilateral(raw, enc, alb, sc, sn, sp, sa):
    raw, alb = raw.astype(np.float64), alb.astype(np.float64)
    n = enc.astype(np.float64) * 2 - 1
    n /= np.linalg.norm(n, axis=2, keepdims=True)
    h, w, _ = raw.shape
    r = max(1, int(np.ceil(3 * np.float32(sp))))
    out = [np.zeros_like(raw) for _ in range(3)]
    for i in range(h):
        for j in range(w):
            ys = np.clip(np.arange(i - r, i + r + 1), 0, h - 1)
            xs = np.clip(np.arange(j - r, j + r + 1), 0, w - 1)
            yy, xx = np.meshgrid(ys, xs, indexing="ij")
            dp = (i - yy) ** 2 + (j - xx) ** 2
            dc = 0.5 * ((raw[i, j] - raw[yy, xx]) ** 2).sum(-1)
            ang = np.arccos(np.clip((n[i, j] * n[yy, xx]).sum(-1), -1, 1))
            da = ((alb[i, j] - alb[yy, xx]) ** 2).sum(-1)
            wg = np.exp(-dp / (2 * sp * sp))
            wc = np.exp(-dp / (2 * sp * sp) - dc / (2 * sc * sc))
            wj = np.exp(-dp / (2 * sp * sp) - dc / (2 * sc * sc) - 1.5 * ang ** 2 / (2 * sn * sn) - da / (2 * sa * sa))
            for k, wt in enumerate((wg, wc, wj)):
                out[k][i, j] = (wt[..., None] * raw[yy, xx]).sum((0, 1)) / wt.sum()
    return out


def _np_atrous(raw, nrm, alb, sc, sn, sa, iters):
    x = raw.astype(np.float64)
    n = nrm.astype(np.float64)
    n = n / np.linalg.norm(n, axis=2, keepdims=True)
    a = alb.astype(np.float64)
    h, w, _ = x.shape
    hk = np.array([1 / 16, 1 / 4, 3 / 8, 1 / 4, 1 / 16])
    lw = np.array([0.2126, 0.7152, 0.0722])
    for it in range(iters):
        s = 1 << it
        lum = x @ lw
        y = np.zeros_like(x)
        for i in range(h):
            for j in range(w):
                acc, ws = np.zeros(3), 0.0
                for dy in range(-2, 3):
                    for dx in range(-2, 3):
                        yi, xj = min(max(i + dy * s, 0), h - 1), min(max(j + dx * s, 0), w - 1)
                        wt = hk[dy + 2] * hk[dx + 2]
                        wt *= np.exp(-(lum[i, j] - lum[yi, xj]) ** 2 / (2 * sc * sc))
                        ang = np.arccos(np.clip(n[i, j] @ n[yi, xj], -1, 1))
                        wt *= np.exp(-ang * ang / (2 * sn * sn))
                        wt *= np.exp(-((a[i, j] - a[yi, xj]) ** 2).sum() / (2 * sa * sa))
                        acc += wt * x[yi, xj]
                        ws += wt
                y[i, j] = acc / ws if ws > 1e-8 else x[i, j]
        x = y
    return x


def _rel(a, b, floor=1e-3):
    return np.abs(a.astype(np.float64) - b) / np.maximum(np.abs(b), floor)


def test_joint_bilateral_matches_float64():
    raw, enc, alb = _inputs(11, 13, 1)
    sig = dict(sigma_c=0.3, sigma_n=0.2, sigma_p=1.2, sigma_a=0.1)
    got = ob.joint_bilateral(raw, enc, alb, **sig)
    want = _np_bilateral(raw, enc, alb, 0.3, 0.2, 1.2, 0.1)
    for g, w_ in zip(got, want):
        assert _rel(g, w_).max() <= 1e-4


def test_bilateral_limits_and_constants():
    raw, enc, alb = _inputs(9, 10, 2)
    g, b, j = ob.joint_bilateral(raw, enc, alb, sigma_c=1e6, sigma_n=1e6, sigma_p=1.5, sigma_a=1e6)
    assert _rel(b, g).max() <= 1e-6 and _rel(j, g).max() <= 1e-6
    const = np.full_like(raw, 0.37)
    for o in ob.joint_bilateral(const, enc, alb, sigma_c=0.3, sigma_n=0.2, sigma_p=2.0, sigma_a=0.1):
        assert np.abs(o - 0.37).max() <= 1e-6


def test_atrous_matches_float64():
    raw, enc, alb = _inputs(12, 14, 3)
    nrm = enc * 2 - 1                                  # the a-trous tool normalises normals as stored
    got = ob.atrous(raw, nrm, alb, sigma_c=0.3, sigma_n=0.2, sigma_a=0.1, iterations=3)
    want = _np_atrous(raw, nrm, alb, 0.3, 0.2, 0.1, 3)
    assert _rel(got, want).max() <= 1e-4


def test_atrous_keeps_a_linear_ramp():
    """With huge range sigmas one pass is the symmetric B3-spline average,
    which keeps a linear function exactly where the 5x5 footprint is inside."""
    h, w = 12, 16
    y, x = np.mgrid[0:h, 0:w]
    ramp = np.repeat((0.05 * x + 0.02 * y + 0.3).astype(np.float32)[..., None], 3, axis=2)
    nrm = np.zeros((h, w, 3), np.float32)
    nrm[..., 2] = 1
    alb = np.full((h, w, 3), 0.5, np.float32)
    out = ob.atrous(ramp, nrm, alb, sigma_c=1e6, sigma_n=1e6, sigma_a=1e6, iterations=1)
    assert np.abs(out[2:-2, 2:-2] - ramp[2:-2, 2:-2]).max() <= 1e-6
    assert np.abs(out[0, 0] - ramp[0, 0]).max() > 1e-3   # clamped border samples bias the corner


@pytest.mark.gpu
@pytest.mark.parametrize("h,w,sp,seed", [(37, 29, 2.0, 1), (64, 80, 0.3, 2), (20, 300, 4.5, 3)])
def test_device_joint_bilateral_matches_oracle(engine_ctx, h, w, sp, seed):
    import rpd_engine as eng
    raw, enc, alb = _inputs(h, w, seed)
    want = ob.joint_bilateral(raw, enc, alb, sigma_c=0.3, sigma_n=0.2, sigma_p=sp, sigma_a=0.1)
    got = eng.joint_bilateral(engine_ctx, raw, enc, alb, sigma_c=0.3, sigma_n=0.2, sigma_p=sp, sigma_a=0.1)
    for g, w_ in zip(got, want):
        print("max rel", _rel(g, w_).max())
        assert _rel(g, w_).max() <= 1e-5


@pytest.mark.gpu
@pytest.mark.parametrize("h,w,iters,seed", [(37, 29, 5, 1), (130, 70, 3, 2), (16, 16, 1, 3)])
def test_device_atrous_matches_oracle(engine_ctx, h, w, iters, seed):
    import rpd_engine as eng
    raw, enc, alb = _inputs(h, w, seed)
    want = ob.atrous(raw, enc, alb, sigma_c=0.3, sigma_n=0.2, sigma_a=0.1, iterations=iters)
    got = eng.atrous(engine_ctx, raw, enc, alb, sigma_c=0.3, sigma_n=0.2, sigma_a=0.1, iterations=iters)
    print("max rel", _rel(got, want).max())
    assert _rel(got, want).max() <= 1e-5
