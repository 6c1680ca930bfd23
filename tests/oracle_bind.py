"""ctypes binding of the CPU oracle (oracle/_build/liboracle.so).

The oracle is a CPU restatement of the reference renderer's hot path; it is
test infrastructure and only tests, the smoke check and the CPU-baseline leg of
bench.py load it.
"""
import ctypes as C
import os
from dataclasses import dataclass, field, fields

import numpy as np

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(REPO, "oracle", "_build", "liboracle.so")


class OrcConfig(C.Structure):
    _fields_ = [
        ("width", C.c_int), ("height", C.c_int), ("spp", C.c_int),
        ("fov", C.c_float), ("focus_distance", C.c_float), ("aperture", C.c_float),
        ("cam_pos", C.c_float * 3), ("cam_dir", C.c_float * 3), ("cam_up", C.c_float * 3),
        ("use_statmc", C.c_int), ("rpf_tile_size", C.c_int), ("color_window_radius", C.c_int),
        ("color_normal_threshold", C.c_float), ("color_depth_threshold", C.c_float),
        ("color_compat_alpha", C.c_float), ("color_sigma_max", C.c_float),
        ("var_window_radius", C.c_int),
        ("var_normal_threshold", C.c_float), ("var_depth_threshold", C.c_float),
        ("var_compat_sigma", C.c_float), ("var_shrinkage_k", C.c_float),
        ("var_iterations", C.c_int), ("adaptive_base_samples", C.c_int), ("adaptive_spp", C.c_int),
        ("adaptive_sigma_max", C.c_float),
        ("adaptive_passes", C.c_int), ("adaptive_importance_smoothing_radius", C.c_int),
        ("rpf_shrinkage_scale", C.c_float), ("rpf_confidence_samples", C.c_float),
        ("max_depth", C.c_int),
    ]


@dataclass
class RenderConfig:
    """Render settings with the reference's defaults (path_tracer/core/config.h:27-48)."""
    width: int = 64
    height: int = 64
    spp: int = 4
    fov: float = 38.6
    focus_distance: float = 1.0
    aperture: float = 0.0
    cam_pos: tuple = (0.278, 0.2744, -0.8)
    cam_dir: tuple = (0.0, 0.0, 1.0)
    cam_up: tuple = (0.0, 1.0, 0.0)
    use_statmc: int = 1
    rpf_tile_size: int = 8
    color_window_radius: int = 7
    color_normal_threshold: float = 0.5
    color_depth_threshold: float = 0.25
    color_compat_alpha: float = 0.05
    color_sigma_max: float = 6.0
    var_window_radius: int = 1
    var_normal_threshold: float = 0.5
    var_depth_threshold: float = 0.25
    var_compat_sigma: float = 30.0
    var_shrinkage_k: float = 1e-3
    var_iterations: int = 2
    adaptive_base_samples: int = 0
    adaptive_spp: int = 4
    adaptive_sigma_max: float = 3.0
    adaptive_passes: int = 3
    adaptive_importance_smoothing_radius: int = 1
    rpf_shrinkage_scale: float = 1.0
    rpf_confidence_samples: float = 128.0
    max_depth: int = 0

    def to_c(self):
        c = OrcConfig()
        for f in fields(self):
            v = getattr(self, f.name)
            if isinstance(v, tuple):
                arr = getattr(c, f.name)
                for i in range(3):
                    arr[i] = v[i]
            else:
                setattr(c, f.name, v)
        return c


def config_c0(width=256, height=256):
    """Plain cornell_box workload: 4 base spp + 3 adaptive passes of 4 spp, defaults."""
    return RenderConfig(width=width, height=height, spp=4, adaptive_spp=4, adaptive_passes=3)


def config_c1(width=1024, height=1024):
    """Full-size cornell_box: 16 base spp + 3x4 adaptive, radius 10, 3 variance
    iterations, thin lens (aperture 0.14, focus 1.05)."""
    return RenderConfig(width=width, height=height, spp=16, adaptive_spp=4, adaptive_passes=3,
                        color_window_radius=10, var_iterations=3, aperture=0.14, focus_distance=1.05)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle library missing: {LIB_PATH} (run make -C oracle)")
        L = C.CDLL(LIB_PATH)
        vp = C.c_void_p
        L.orc_job_create.restype = vp
        L.orc_job_create.argtypes = [C.POINTER(OrcConfig), C.c_uint32, C.c_int]
        L.orc_job_create_scene.restype = vp
        L.orc_job_create_scene.argtypes = [C.POINTER(OrcConfig), C.c_uint32, C.c_int, C.c_void_p, C.c_int,
                                           C.c_void_p, C.c_int, C.c_void_p, C.c_int, C.c_void_p, C.c_int,
                                           C.c_void_p]
        for name in ("orc_job_destroy", "orc_render", "orc_recompute_sensitivity", "orc_statmc",
                     "orc_variance", "orc_finalize"):
            getattr(L, name).argtypes = [vp]
            getattr(L, name).restype = None
        L.orc_job_flags.argtypes = [vp, C.c_int, C.c_int, C.c_int]
        L.orc_pass.argtypes = [vp, C.c_void_p, C.c_int]
        L.orc_adaptive_counts.argtypes = [vp, C.c_void_p]
        L.orc_adaptive_counts_from_importance.argtypes = [vp, C.c_void_p, C.c_void_p]
        L.orc_stage_ms.argtypes = [vp, C.c_int]
        L.orc_stage_ms.restype = C.c_double
        L.orc_dims.argtypes = [vp, C.c_void_p]
        L.orc_get.argtypes = [vp, C.c_char_p, C.c_void_p]
        L.orc_get.restype = C.c_long
        L.orc_set.argtypes = [vp, C.c_char_p, C.c_void_p]
        L.orc_set.restype = C.c_long
        L.orc_accumulate_records.argtypes = [vp, C.c_void_p, C.c_void_p, C.c_long]
        L.orc_num_records.argtypes = [vp]
        L.orc_num_records.restype = C.c_long
        L.orc_hash_u32.argtypes = [C.c_uint32]
        L.orc_hash_u32.restype = C.c_uint32
        L.orc_hash_to_unit_float.argtypes = [C.c_uint32]
        L.orc_hash_to_unit_float.restype = C.c_float
        L.orc_pixel_sample_seed.argtypes = [C.c_uint32, C.c_int, C.c_int]
        L.orc_pixel_sample_seed.restype = C.c_uint32
        L.orc_sampler_draws.argtypes = [C.c_uint32, C.c_int, C.c_void_p]
        L.orc_sampler_pairs.argtypes = [C.c_uint32, C.c_int, C.c_void_p]
        L.orc_mt_u32.argtypes = [C.c_uint32, C.c_int]
        L.orc_mt_u32.restype = C.c_uint32
        L.orc_t_critical.argtypes = [C.c_float, C.c_float]
        L.orc_t_critical.restype = C.c_float
        L.orc_sample_disk.argtypes = [C.c_float, C.c_float, C.c_void_p]
        L.orc_sample_sphere.argtypes = [C.c_float, C.c_float, C.c_void_p]
        L.orc_accumulate_samples.argtypes = [C.c_void_p, C.c_int, C.c_void_p]
        L.orc_binned.argtypes = [C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_void_p]
        L.orc_dep2d.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int]
        L.orc_dep2d.restype = C.c_float
        L.orc_confidence.argtypes = [C.c_double, C.c_float]
        L.orc_confidence.restype = C.c_float
        L.orc_relative_random_sensitivity.argtypes = [C.c_float, C.c_float]
        L.orc_relative_random_sensitivity.restype = C.c_float
        L.orc_shared_reliability.argtypes = [C.c_void_p, C.c_void_p]
        L.orc_shared_reliability.restype = C.c_float
        L.orc_relax_compatibility.argtypes = [C.c_double, C.c_float, C.c_float]
        L.orc_relax_compatibility.restype = C.c_double
        L.orc_conservative_variance_update.argtypes = [C.c_float, C.c_float, C.c_float]
        L.orc_conservative_variance_update.restype = C.c_float
        L.orc_brdf_sample.argtypes = [C.c_int, C.c_void_p, C.c_float, C.c_void_p, C.c_void_p,
                                      C.c_float, C.c_float, C.c_void_p]
        L.orc_brdf_eval.argtypes = [C.c_int, C.c_void_p, C.c_float, C.c_void_p, C.c_void_p,
                                    C.c_void_p, C.c_void_p]
        L.orc_cornell_triangles.argtypes = [C.c_void_p]
        L.orc_cornell_triangles.restype = C.c_long
        L.orc_cornell_hits.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p]
        L.orc_image_metrics.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_void_p]
        L.orc_joint_bilateral.argtypes = [vp, vp, vp, C.c_int, C.c_int, C.c_float, C.c_float, C.c_float, C.c_float,
                                          vp, vp, vp]
        L.orc_atrous.argtypes = [vp, vp, vp, C.c_int, C.c_int, C.c_float, C.c_float, C.c_float, C.c_int, vp]
        _lib = L
    return _lib


def ptr(a):
    return a.ctypes.data_as(C.c_void_p)


METRIC_NAMES = ("mae", "rmse", "psnr", "peak", "ssim_r", "ssim_g", "ssim_b", "ssim_mean", "nrmse", "log_ssim11",
                "log_gradient_nrmse", "relmse")


def _rgb(a):
    a = np.ascontiguousarray(a, np.float32)
    assert a.ndim == 3 and a.shape[2] == 3
    return a


def joint_bilateral(raw, normal, albedo, sigma_c=1.0, sigma_n=1.0, sigma_p=1.0, sigma_a=1.0):
    """denoising/joint_bilateral.cpp filter(): (gaussian, bilateral, joint), each (H, W, 3)."""
    raw, normal, albedo = _rgb(raw), _rgb(normal), _rgb(albedo)
    h, w = raw.shape[:2]
    outs = [np.zeros_like(raw) for _ in range(3)]
    lib().orc_joint_bilateral(ptr(raw), ptr(normal), ptr(albedo), w, h, sigma_c, sigma_n, sigma_p, sigma_a,
                              *[ptr(o) for o in outs])
    return tuple(outs)


def atrous(raw, normal, albedo, sigma_c=0.5, sigma_n=0.5, sigma_a=0.1, iterations=5):
    """denoising/atrous_wavelet.cpp filter(): the filtered (H, W, 3) image."""
    raw, normal, albedo = _rgb(raw), _rgb(normal), _rgb(albedo)
    h, w = raw.shape[:2]
    out = np.zeros_like(raw)
    lib().orc_atrous(ptr(raw), ptr(normal), ptr(albedo), w, h, sigma_c, sigma_n, sigma_a, iterations, ptr(out))
    return out


def image_metrics(ref, test):
    """eval_tools.cpp hdr_metrics + compute_extended_metrics (+ relMSE) on
    (H, W, 3) or (N, 3) float images of equal shape, given the width."""
    ref = np.ascontiguousarray(ref, np.float32)
    test = np.ascontiguousarray(test, np.float32)
    assert ref.shape == test.shape and ref.ndim == 3 and ref.shape[2] == 3
    h, w = ref.shape[:2]
    out = np.zeros(len(METRIC_NAMES), np.float64)
    lib().orc_image_metrics(ptr(ref), ptr(test), w, h, ptr(out))
    return dict(zip(METRIC_NAMES, out.tolist()))


# shapes of the per-job buffers: name -> (dtype, per-pixel width)
_PIXEL_BUFFERS = {
    "radiance": (np.float32, 3), "raw_radiance": (np.float32, 3), "albedo": (np.float32, 3),
    "normal": (np.float32, 3), "world_pos": (np.float32, 3), "denoised": (np.float32, 3),
    "depth": (np.float32, 1), "uncertainty": (np.float32, 1), "alpha": (np.float32, 1),
    "variance_denoised": (np.float32, 1), "var_total": (np.float32, 1), "var_eff": (np.float32, 1),
    "var_ratio": (np.float32, 1), "sensitivity_confidence": (np.float32, 1),
    "sensitivity_gradient": (np.float32, 1), "adaptive_importance": (np.float32, 1),
    "hit_count": (np.int32, 1), "sample_count": (np.int32, 1), "stats": (np.float32, 12),
    "features": (np.float32, 7),
}


class OracleJob:
    """One render job of the CPU oracle (mirrors the reference's process state)."""

    def __init__(self, cfg: RenderConfig, seed: int, nthreads: int = 1, skip_dead=False,
                 keep_records=False, verbose=False, scene=None):
        """scene: None for the built-in cornell_box, or (triangles, spheres, materials,
        lights, env, counts) ctypes arrays in the engine's scene layout."""
        self.L = lib()
        self.cfg = cfg
        self._c = cfg.to_c()
        if scene is None:
            self.h = self.L.orc_job_create(C.byref(self._c), seed & 0xFFFFFFFF, nthreads)
        else:
            tris, sph, mats, lights, env, n = scene
            self.h = self.L.orc_job_create_scene(C.byref(self._c), seed & 0xFFFFFFFF, nthreads, tris, n[0], sph,
                                                 n[1], mats, n[2], lights, n[3], env)
        self.L.orc_job_flags(self.h, int(skip_dead), int(keep_records), int(verbose))
        d = np.zeros(6, np.int32)
        self.L.orc_dims(self.h, ptr(d))
        self.w, self.ht, self.nodes_x, self.nodes_y, self.step, self.support = (int(v) for v in d)

    def __del__(self):
        if getattr(self, "h", None):
            self.L.orc_job_destroy(self.h)
            self.h = None

    def render(self):
        self.L.orc_render(self.h)

    def run_pass(self, counts, statmc=True):
        counts = np.ascontiguousarray(counts, np.int32)
        assert counts.size == self.w * self.ht
        self.L.orc_pass(self.h, ptr(counts), int(statmc))

    def accumulate_records(self, counts, records32):
        """Fold given records (32-float rows, pixel-major) like one StatMC pass."""
        counts = np.ascontiguousarray(counts, np.int32)
        rec = np.ascontiguousarray(records32, np.float32)
        assert counts.size == self.w * self.ht and rec.shape[0] == int(counts.sum())
        self.L.orc_accumulate_records(self.h, ptr(counts), ptr(rec), rec.shape[0])

    def recompute_sensitivity(self):
        self.L.orc_recompute_sensitivity(self.h)

    def statmc(self):
        self.L.orc_statmc(self.h)

    def variance(self):
        self.L.orc_variance(self.h)

    def adaptive_counts(self):
        out = np.zeros(self.w * self.ht, np.int32)
        self.L.orc_adaptive_counts(self.h, ptr(out))
        return out

    def counts_from_importance(self, importance):
        imp = np.ascontiguousarray(importance, np.float32)
        out = np.zeros(self.w * self.ht, np.int32)
        self.L.orc_adaptive_counts_from_importance(self.h, ptr(imp), ptr(out))
        return out

    def finalize(self):
        self.L.orc_finalize(self.h)

    def stage_ms(self, i):
        return self.L.orc_stage_ms(self.h, i)

    def get(self, name):
        n = self.w * self.ht
        if name == "grid":
            out = np.zeros(self.nodes_x * self.nodes_y * 440, np.float64)
        elif name == "tiles":
            out = np.zeros(18 * self.nodes_x * self.nodes_y, np.float32)
        elif name == "records":
            out = np.zeros(max(1, self.L.orc_num_records(self.h)) * 32, np.float32)
        else:
            dt, k = _PIXEL_BUFFERS[name]
            out = np.zeros(n * k, dt)
        got = self.L.orc_get(self.h, name.encode(), ptr(out))
        if got < 0:
            raise KeyError(name)
        out = out[:got]
        if name == "grid":
            return out.reshape(-1, 440)
        if name == "tiles":
            return out.reshape(18, -1)
        if name == "records":
            return out.reshape(-1, 32)
        k = _PIXEL_BUFFERS[name][1]
        return out.reshape(n, k) if k > 1 else out

    def set(self, name, arr):
        if name == "grid":
            arr = np.ascontiguousarray(arr, np.float64)
        else:
            arr = np.ascontiguousarray(arr, _PIXEL_BUFFERS[name][0])
        got = self.L.orc_set(self.h, name.encode(), ptr(arr))
        if got < 0:
            raise KeyError(name)
