"""rpd_image_metrics_compute_noalloc: the allocation-free variant of
rpd_image_metrics_compute for device images. Its partial sums live in the
library's fixed device array, so it must give the same metrics as the
allocating entry point, call after call, for a smaller image after a larger
one (the array keeps the last call's contents), and refuse images beyond its
stated maximum. Device images are allocated by the test through the HIP
runtime; the entry point itself allocates nothing."""
import ctypes as C

import numpy as np
import pytest

import rpd_engine as eng

pytestmark = pytest.mark.gpu

HIP = None


def _hip():
    global HIP
    if HIP is None:
        HIP = C.CDLL("libamdhip64.so")
        HIP.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        HIP.hipFree.argtypes = [C.c_void_p]
        HIP.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    return HIP


class DeviceImage:
    """An (H, W, 3) float32 image copied to device memory."""

    def __init__(self, host):
        host = np.ascontiguousarray(host, np.float32)
        self.p = C.c_void_p()
        assert _hip().hipMalloc(C.byref(self.p), host.nbytes) == 0
        assert _hip().hipMemcpy(self.p, host.ctypes.data_as(C.c_void_p), host.nbytes, 1) == 0   # host to device

    def free(self):
        if self.p:
            _hip().hipFree(self.p)
            self.p = C.c_void_p()


def _images(rng, h, w):
    ref = rng.lognormal(-1.0, 1.0, (h, w, 3)).astype(np.float32)
    test = (ref * rng.lognormal(0.0, 0.3, (h, w, 3))).astype(np.float32)
    return ref, test


def _metrics(ctx, fn, ref, test, h, w):
    m = eng.ImageMetrics()
    if fn == "noalloc":
        rc = eng.lib().rpd_image_metrics_compute_noalloc(ctx.h, ref.p, test.p, w, h, C.byref(m))
    else:
        rc = eng.lib().rpd_image_metrics_compute(ctx.h, ref.p, test.p, w, h, 1, C.byref(m))
    eng.check(rc)
    return m.as_dict()


def test_noalloc_metrics_equal_the_allocating_entry_point():
    ctx = eng.Context(0)
    rng = np.random.default_rng(3)
    held = []
    try:
        for h, w in [(700, 1100), (37, 23)]:          # the second call smaller than the first
            ref_h, test_h = _images(rng, h, w)
            ref, test = DeviceImage(ref_h), DeviceImage(test_h)
            held += [ref, test]
            print(f"noalloc metrics {w}x{h}", flush=True)
            a = _metrics(ctx, "noalloc", ref, test, h, w)
            b = _metrics(ctx, "alloc", ref, test, h, w)
            assert a == b, (w, h, a, b)
            assert 0.0 < a["relmse"] and 0.0 < a["nrmse"] and a["log_ssim11"] < 1.0
    finally:
        for d in held:
            d.free()
        ctx.close()


def test_noalloc_metrics_refuse_images_beyond_the_maximum():
    ctx = eng.Context(0)
    h, w = 4096, 4097                                  # 256 x 257 tiles, one column of tiles too many
    img = np.zeros((h, w, 3), np.float32)
    ref, test = DeviceImage(img), DeviceImage(img)
    try:
        m = eng.ImageMetrics()
        rc = eng.lib().rpd_image_metrics_compute_noalloc(ctx.h, ref.p, test.p, w, h, C.byref(m))
        assert rc == -1
        assert b"4096 x 4096" in eng.lib().rpd_last_error()
        # the largest image it takes
        a = _metrics(ctx, "noalloc", ref, test, 4096, 4096)
        assert a["mae"] == 0.0 and a["relmse"] == 0.0
    finally:
        ref.free()
        test.free()
        ctx.close()
