"""ctypes binding of the engine's C ABI (include/rpd_engine.h).

Loads lib/librpd_engine.so next to this file and raises if it is missing: there
is no fallback path. Mirrors the reference renderer's stage functions
(path_tracer/main.cpp: render, recompute_sensitivity, statmc_denoise,
variance_denoise, compute_adaptive_sample_counts) on one job's device state.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "librpd_engine.so")

# buffer ids (include/rpd_engine.h)
BUF = dict(
    radiance=0, raw_radiance=1, albedo=2, normal=3, world_pos=4, denoised=5,
    depth=10, uncertainty=11, alpha=12, variance_denoised=13, var_total=14, var_eff=15, var_ratio=16,
    sensitivity_confidence=17, sensitivity_gradient=18, adaptive_importance=19, sample_fraction=20,
    sens_all=30, sens_pixel=31, sens_brdf=32, sens_lens=33, sens_light=34, sens_light_uv=35,
    sens_light_select=36, sens_environment=37, sens_rr=38,
    sample_count=50, hit_count=51, pass_counts=52,
    stats=60, sums=61, grid=62, tiles=63, features=64, records=65,
    debug_neighbors=70, debug_weight_sum=71, debug_neff=72, debug_variance_source=73,
)
_RGB = {"radiance", "raw_radiance", "albedo", "normal", "world_pos", "denoised"}
_INT = {"sample_count", "hit_count", "pass_counts"}
_WIDTH = {"stats": 12, "sums": 13, "features": 7, "records": 32}


class RenderConfig(C.Structure):
    """POD mirror of RenderConfig (path_tracer/core/config.h:15-49)."""
    _fields_ = [
        ("image_width", C.c_int), ("image_height", C.c_int), ("samples_per_pixel", C.c_int),
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


class Camera(C.Structure):
    _fields_ = [("position", C.c_float * 3), ("direction", C.c_float * 3), ("world_up", C.c_float * 3),
                ("fov_degrees", C.c_float), ("focus_distance", C.c_float), ("aperture", C.c_float)]


class SynthConfig(C.Structure):
    """rpd_synth_config: synthetic per-sample records of the denoise-only workload."""
    _fields_ = [("seed", C.c_uint32), ("samples_per_pixel", C.c_int), ("region_size", C.c_int),
                ("log_mean", C.c_float), ("log_sigma", C.c_float), ("firefly_fraction", C.c_float),
                ("firefly_scale", C.c_float), ("correlation", C.c_float)]


class ProceduralConfig(C.Structure):
    """rpd_procedural_config: the icosphere + spheres benchmark scene."""
    _fields_ = [("seed", C.c_uint32), ("subdivision_level", C.c_int), ("n_spheres", C.c_int),
                ("displacement", C.c_float), ("noise_frequency", C.c_float), ("noise_octaves", C.c_int),
                ("environment", C.c_float * 3), ("light_emission", C.c_float)]


class Triangle(C.Structure):
    _fields_ = [("v0", C.c_float * 3), ("v1", C.c_float * 3), ("v2", C.c_float * 3), ("normal", C.c_float * 3),
                ("emission", C.c_float * 3), ("material_id", C.c_int), ("light_id", C.c_int)]


class Sphere(C.Structure):
    _fields_ = [("center", C.c_float * 3), ("radius", C.c_float), ("emission", C.c_float * 3),
                ("material_id", C.c_int), ("light_id", C.c_int)]


class Material(C.Structure):
    _fields_ = [("kind", C.c_int), ("albedo", C.c_float * 3), ("roughness", C.c_float), ("n1", C.c_float * 3),
                ("n2", C.c_float * 3), ("distribution", C.c_int)]


class Light(C.Structure):
    _fields_ = [("kind", C.c_int), ("p0", C.c_float * 3), ("p1", C.c_float * 3), ("p2", C.c_float * 3),
                ("normal", C.c_float * 3), ("emission", C.c_float * 3)]


class ServerConfig(C.Structure):
    _fields_ = [("max_in_flight", C.c_int), ("memory_budget", C.c_uint64), ("graph_passes", C.c_int),
                ("cobatch_tail", C.c_int)]


class ServerStats(C.Structure):
    _fields_ = [("submitted", C.c_uint64), ("completed", C.c_uint64), ("failed", C.c_uint64),
                ("max_in_flight_seen", C.c_int), ("peak_bytes", C.c_uint64),
                ("latency_p50_ms", C.c_double), ("latency_p99_ms", C.c_double), ("latency_mean_ms", C.c_double),
                ("latency_max_ms", C.c_double), ("queue_p50_ms", C.c_double), ("queue_p99_ms", C.c_double),
                ("tail_batches", C.c_uint64), ("tail_jobs", C.c_uint64), ("over_budget", C.c_uint64)]


class ImageMetrics(C.Structure):
    """rpd_image_metrics (shared/tools/eval_tools.cpp metrics + relMSE)."""
    _fields_ = [("mae", C.c_double), ("rmse", C.c_double), ("psnr", C.c_double), ("peak", C.c_double),
                ("ssim", C.c_double * 3), ("ssim_mean", C.c_double), ("nrmse", C.c_double),
                ("log_ssim11", C.c_double), ("log_gradient_nrmse", C.c_double), ("relmse", C.c_double)]

    def as_dict(self):
        d = {k: getattr(self, k) for k, _ in self._fields_ if k != "ssim"}
        d.update(ssim_r=self.ssim[0], ssim_g=self.ssim[1], ssim_b=self.ssim[2])
        return d


class BilateralConfig(C.Structure):
    _fields_ = [("sigma_c", C.c_float), ("sigma_n", C.c_float), ("sigma_p", C.c_float), ("sigma_a", C.c_float)]


class AtrousConfig(C.Structure):
    _fields_ = [("sigma_c", C.c_float), ("sigma_n", C.c_float), ("sigma_a", C.c_float), ("iterations", C.c_int)]


# rpd_sample_record as a numpy structured type (100 bytes)
SAMPLE_RECORD = np.dtype([
    ("radiance", np.float32, 3), ("flags", np.uint32), ("normal", np.float32, 3), ("depth", np.float32),
    ("albedo", np.float32, 3), ("position", np.float32, 3), ("pixel_u", np.float32, 2), ("lens_u", np.float32, 2),
    ("brdf_u", np.float32, 2), ("light_u", np.float32, 2), ("env_u", np.float32, 2), ("light_index", np.int32)])
assert SAMPLE_RECORD.itemsize == 100


def records_from_rows(rows):
    """rpd_sample_record array from 32-float record rows (the RPD_BUF_RECORDS layout:
    pixel, rgb, finite, hit, normal, albedo, position, depth, pixel_u, lens_u, brdf_u,
    light_u, env_u, light index, rr survived, validity bits)."""
    rows = np.asarray(rows, np.float32)
    r = np.zeros(rows.shape[0], SAMPLE_RECORD)
    r["radiance"] = rows[:, 1:4]
    r["normal"] = rows[:, 6:9]
    r["albedo"] = rows[:, 9:12]
    r["position"] = rows[:, 12:15]
    r["depth"] = rows[:, 15]
    r["pixel_u"] = rows[:, 16:18]
    r["lens_u"] = rows[:, 18:20]
    r["brdf_u"] = rows[:, 20:22]
    r["light_u"] = rows[:, 22:24]
    r["env_u"] = rows[:, 24:26]
    r["light_index"] = np.maximum(rows[:, 26], 0).astype(np.int32)
    v = rows[:, 28].astype(np.uint32)
    fl = np.where(rows[:, 5] != 0, 2, 0).astype(np.uint32)
    fl |= np.where(v & 1, 4, 0).astype(np.uint32)        # pixel
    fl |= np.where(v & 2, 8, 0).astype(np.uint32)        # lens
    fl |= np.where(v & 4, 16, 0).astype(np.uint32)       # light select
    fl |= np.where(v & 8, 32, 0).astype(np.uint32)       # brdf
    fl |= np.where(v & 16, 64, 0).astype(np.uint32)      # light uv
    fl |= np.where(v & 32, 128, 0).astype(np.uint32)     # environment
    fl |= np.where(v & 64, 256, 0).astype(np.uint32)     # russian roulette
    fl |= np.where(rows[:, 27] != 0, 512, 0).astype(np.uint32)
    r["flags"] = fl
    return r


# scene POD enums (include/rpd_engine.h)
RPD_MAT_LAMBERTIAN, RPD_MAT_MICROFACET, RPD_MAT_MIRROR = 0, 1, 2
RPD_LIGHT_RECT, RPD_LIGHT_AREA, RPD_LIGHT_POINT, RPD_LIGHT_DIRECTIONAL = 0, 1, 2, 3


class EngineError(RuntimeError):
    pass


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise EngineError(f"engine library not built: {LIB_PATH}")
        L = C.CDLL(LIB_PATH)
        vp = C.c_void_p
        L.rpd_last_error.restype = C.c_char_p
        L.rpd_render_config_defaults.argtypes = [C.POINTER(RenderConfig)]
        L.rpd_render_config_defaults.restype = None
        L.rpd_procedural_config_defaults.argtypes = [C.POINTER(ProceduralConfig)]
        L.rpd_procedural_config_defaults.restype = None
        L.rpd_bilateral_config_defaults.argtypes = [C.POINTER(BilateralConfig)]
        L.rpd_bilateral_config_defaults.restype = None
        L.rpd_atrous_config_defaults.argtypes = [C.POINTER(AtrousConfig)]
        L.rpd_atrous_config_defaults.restype = None
        L.rpd_synth_config_defaults.argtypes = [C.POINTER(SynthConfig)]
        L.rpd_synth_config_defaults.restype = None
        for name, args in [
            ("rpd_ctx_create", [C.c_int, C.POINTER(vp)]), ("rpd_ctx_destroy", [vp]),
            ("rpd_ctx_set_memory_budget", [vp, C.c_uint64]),
            ("rpd_ctx_set_bvh_builder", [vp, C.c_int]),
            ("rpd_ctx_set_buffer_cache", [vp, C.c_uint64, C.POINTER(C.c_uint64)]),
            ("rpd_ctx_set_held_limit", [vp, C.c_uint64]),
            ("rpd_ctx_set_tape_prefix", [vp, C.c_int, C.c_int]),
            ("rpd_scene_create", [vp, vp, C.c_int, vp, C.c_int, vp, C.c_int, vp, C.c_int, vp, C.POINTER(vp)]),
            ("rpd_scene_builtin", [vp, C.c_char_p, C.POINTER(vp)]),
            ("rpd_scene_counts", [vp, vp, vp, vp, vp]),
            ("rpd_scene_destroy", [vp]),
            ("rpd_job_create", [vp, vp, C.POINTER(RenderConfig), C.POINTER(Camera), C.c_uint32, C.POINTER(vp)]),
            ("rpd_job_submit", [vp, vp, C.POINTER(RenderConfig), C.POINTER(Camera), C.c_uint32, C.POINTER(vp)]),
            ("rpd_job_poll", [vp, C.POINTER(C.c_int)]), ("rpd_job_wait", [vp]), ("rpd_job_release", [vp]),
            ("rpd_job_run_pass", [vp, vp, C.c_int]), ("rpd_job_run_adaptive_pass", [vp]),
            ("rpd_job_recompute_sensitivity", [vp]), ("rpd_job_statmc_denoise", [vp]),
            ("rpd_job_variance_denoise", [vp]), ("rpd_job_adaptive_counts", [vp, vp]),
            ("rpd_job_adaptive_counts_from_importance", [vp, vp, vp]),
            ("rpd_job_finalize", [vp]), ("rpd_job_sync", [vp]),
            ("rpd_job_buffer_size", [vp, C.c_int, C.POINTER(C.c_size_t)]),
            ("rpd_job_read", [vp, C.c_int, vp, C.c_size_t]), ("rpd_job_write", [vp, C.c_int, vp, C.c_size_t]),
            ("rpd_job_grid_dims", [vp, vp]), ("rpd_job_stage_ms", [vp, vp, C.c_int]),
            ("rpd_job_counters", [vp, vp, C.c_int]),
            ("rpd_job_create_denoiser", [vp, C.POINTER(RenderConfig), C.c_int, C.POINTER(vp)]),
            ("rpd_job_load_records", [vp, vp, vp, C.c_size_t]),
            ("rpd_job_synthesize_records", [vp, C.POINTER(SynthConfig)]),
            ("rpd_job_denoise_records", [vp]),
            ("rpd_denoise_batch", [vp, vp, C.c_int, C.POINTER(C.c_double)]),
            ("rpd_job_submit_passes", [vp, vp, C.POINTER(RenderConfig), C.POINTER(Camera), C.c_uint32, C.POINTER(vp)]),
            ("rpd_render_finish_batch", [vp, vp, C.c_int, C.POINTER(C.c_double)]),
            ("rpd_ctx_set_graph_passes", [vp, C.c_int]),
            ("rpd_ctx_set_trace_wavefront", [vp, C.c_float]),
            ("rpd_ctx_set_gather_overlap", [vp, C.c_int]),
            ("rpd_ctx_set_statmc_pairs", [vp, C.c_int]),
            ("rpd_ctx_set_debug_outputs", [vp, C.c_int]),
            ("rpd_statmc_tile_shapes", [vp, C.c_int, C.POINTER(C.c_int)]),
            ("rpd_ctx_set_statmc_tile", [vp, C.c_int, C.c_int]),
            ("rpd_ctx_statmc_tile", [vp, C.c_int, C.c_int64, C.POINTER(C.c_int), C.POINTER(C.c_int)]),
            ("rpd_abi_struct_sizes", [vp, C.c_int]),
            ("rpd_image_metrics_compute", [vp, vp, vp, C.c_int, C.c_int, C.c_int, C.POINTER(ImageMetrics)]),
            ("rpd_image_metrics_compute_noalloc", [vp, vp, vp, C.c_int, C.c_int, C.POINTER(ImageMetrics)]),
            ("rpd_job_image_metrics", [vp, C.c_int, vp, C.c_int, C.POINTER(ImageMetrics)]),
            ("rpd_builtin_scene_arrays", [C.c_char_p, vp, vp, vp, vp, vp, vp, vp, vp, vp]),
            ("rpd_obj_load", [C.c_char_p, C.c_int, vp, vp]),
            ("rpd_joint_bilateral", [vp, vp, vp, vp, C.c_int, C.c_int, C.POINTER(BilateralConfig), vp, vp, vp, C.c_int]),
            ("rpd_atrous", [vp, vp, vp, vp, C.c_int, C.c_int, C.POINTER(AtrousConfig), vp, C.c_int]),
            ("rpd_procedural_scene_arrays", [C.POINTER(ProceduralConfig), vp, vp, vp, vp, vp, vp, vp, vp, vp]),
            ("rpd_scene_procedural", [vp, C.POINTER(ProceduralConfig), C.POINTER(vp)]),
            ("rpd_scene_query", [vp, C.c_int, vp, vp, vp, vp, C.c_int, vp, vp, vp, vp]),
            ("rpd_scene_bvh_info", [vp, vp, vp]),
            ("rpd_scene_bvh_builder", [vp, vp]),
            ("rpd_job_bytes", [C.POINTER(RenderConfig), C.POINTER(C.c_uint64)]),
            ("rpd_ctx_memory", [vp, vp, vp, vp, vp]),
            ("rpd_server_create", [vp, C.POINTER(ServerConfig), C.POINTER(vp)]),
            ("rpd_server_submit", [vp, vp, C.POINTER(RenderConfig), C.POINTER(Camera), C.c_uint32,
                                   C.POINTER(C.c_uint64)]),
            ("rpd_server_wait", [vp, C.c_uint64, C.POINTER(vp), C.POINTER(C.c_double)]),
            ("rpd_server_stats_get", [vp, C.POINTER(ServerStats)]),
            ("rpd_server_destroy", [vp]),
        ]:
            f = getattr(L, name)
            f.argtypes = args
            f.restype = C.c_int
        _lib = L
    return _lib


def check(rc):
    if rc != 0:
        raise EngineError(f"engine error {rc}: {lib().rpd_last_error().decode(errors='replace')}")


def default_config(**kw):
    c = RenderConfig()
    lib().rpd_render_config_defaults(C.byref(c))
    for k, v in kw.items():
        setattr(c, k, v)
    return c


def make_camera(position=(0.278, 0.2744, -0.8), direction=(0.0, 0.0, 1.0), world_up=(0.0, 1.0, 0.0),
                fov=38.6, focus_distance=1.0, aperture=0.0):
    cam = Camera()
    for i in range(3):
        cam.position[i] = position[i]
        cam.direction[i] = direction[i]
        cam.world_up[i] = world_up[i]
    cam.fov_degrees = fov
    cam.focus_distance = focus_distance
    cam.aperture = aperture
    return cam


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p)


class Context:
    def __init__(self, device=0):
        self.L = lib()
        h = C.c_void_p()
        check(self.L.rpd_ctx_create(device, C.byref(h)))
        self.h = h
        self._scenes = {}

    def scene(self, name="cornell_box"):
        if name not in self._scenes:
            s = C.c_void_p()
            check(self.L.rpd_scene_builtin(self.h, name.encode(), C.byref(s)))
            self._scenes[name] = s
        return self._scenes[name]

    def procedural_scene(self, name, pc):
        """Creates (once) the procedural scene `pc` on the device under `name`."""
        if name not in self._scenes:
            s = C.c_void_p()
            check(self.L.rpd_scene_procedural(self.h, C.byref(pc), C.byref(s)))
            self._scenes[name] = s
        return self._scenes[name]

    def set_statmc_tile(self, tile_w, tile_h):
        """Fix the StatMC window tile shape; (0, 0) returns to tuning."""
        check(self.L.rpd_ctx_set_statmc_tile(self.h, int(tile_w), int(tile_h)))

    def statmc_tile(self, radius, pixels):
        """The tile shape used for this radius and pixel count, (0, 0) if not tuned yet."""
        w, h = C.c_int(), C.c_int()
        check(self.L.rpd_ctx_statmc_tile(self.h, int(radius), int(pixels), C.byref(w), C.byref(h)))
        return w.value, h.value

    def set_debug_outputs(self, on):
        """Jobs created next keep the debug_statmc_outputs buffers (debug_*)."""
        check(self.L.rpd_ctx_set_debug_outputs(self.h, int(on)))

    def set_graph_passes(self, on):
        check(self.L.rpd_ctx_set_graph_passes(self.h, int(on)))

    def set_statmc_pairs(self, on):
        """StatMC window as the symmetric pair and sum passes (on, the default) or one pass; same bits."""
        check(self.L.rpd_ctx_set_statmc_pairs(self.h, int(on)))

    def set_gather_overlap(self, on):
        """Jobs created next gather a pass into the RPF grid beside the next
        pass's trace (on, the default) or in line (off); same outputs."""
        check(self.L.rpd_ctx_set_gather_overlap(self.h, int(on)))

    def set_trace_wavefront(self, rounds_per_sample):
        """Wavefront trace for flat scenes: rounds per sample of a pass before the path-regeneration
        kernel finishes it (0: the path-regeneration kernel alone)."""
        check(self.L.rpd_ctx_set_trace_wavefront(self.h, float(rounds_per_sample)))

    def memory(self):
        v = [C.c_uint64() for _ in range(4)]
        check(self.L.rpd_ctx_memory(self.h, *[C.byref(x) for x in v]))
        return dict(in_use=v[0].value, budget=v[1].value, device_free=v[2].value, device_total=v[3].value)

    def set_memory_budget(self, nbytes):
        check(self.L.rpd_ctx_set_memory_budget(self.h, int(nbytes)))

    def set_bvh_builder(self, builder):
        """RPD_BVH_HOST_SAH or RPD_BVH_DEVICE_LBVH for the scenes created next."""
        check(self.L.rpd_ctx_set_bvh_builder(self.h, int(builder)))

    def scene_query(self, name, origins, dirs, t_max, occlusion, use_bvh):
        """Host-side queries of scene `name` (its BVH as built, or the exhaustive loop)."""
        return _scene_query(self.L, self.scene(name), origins, dirs, t_max, occlusion, use_bvh)

    def bvh_info(self, name):
        return _bvh_info(self.L, self.scene(name))

    def bvh_builder(self, name):
        b = C.c_int()
        check(self.L.rpd_scene_bvh_builder(self.scene(name), C.byref(b)))
        return b.value

    def set_buffer_cache(self, nbytes):
        """Bytes of released jobs' device slabs kept for reuse (0 disables); returns the bytes now kept."""
        kept = C.c_uint64()
        check(self.L.rpd_ctx_set_buffer_cache(self.h, int(nbytes), C.byref(kept)))
        return kept.value

    def set_tape_prefix(self, words_per_sample, margin):
        """Tape words written before a pass: per sample of the pixel, and beyond (defaults 48, 64)."""
        check(self.L.rpd_ctx_set_tape_prefix(self.h, int(words_per_sample), int(margin)))

    def set_held_limit(self, nbytes):
        """Held bytes up to which a job may take a kept slab larger than it needs (0: no limit)."""
        check(self.L.rpd_ctx_set_held_limit(self.h, int(nbytes)))

    def close(self):
        for s in self._scenes.values():
            self.L.rpd_scene_destroy(s)
        self._scenes = {}
        if self.h:
            self.L.rpd_ctx_destroy(self.h)
            self.h = None


def image_metrics(ctx, ref, test):
    """Device metrics of host image `test` against `ref`, both (H, W, 3) float32."""
    ref = np.ascontiguousarray(ref, np.float32)
    test = np.ascontiguousarray(test, np.float32)
    if ref.shape != test.shape or ref.ndim != 3 or ref.shape[2] != 3:
        raise ValueError("images must both be (H, W, 3)")
    m = ImageMetrics()
    check(lib().rpd_image_metrics_compute(ctx.h, _ptr(ref), _ptr(test), ref.shape[1], ref.shape[0], 0, C.byref(m)))
    return m.as_dict()


def _rgb_image(a):
    a = np.ascontiguousarray(a, np.float32)
    if a.ndim != 3 or a.shape[2] != 3:
        raise ValueError("images must be (H, W, 3)")
    return a


def joint_bilateral(ctx, raw, normal, albedo, **sigmas):
    """denoising/joint_bilateral.cpp on the device: (gaussian, bilateral, joint)
    images from host images (normal encoded as (n + 1) / 2)."""
    raw, normal, albedo = _rgb_image(raw), _rgb_image(normal), _rgb_image(albedo)
    cfg = BilateralConfig()
    lib().rpd_bilateral_config_defaults(C.byref(cfg))
    for k, v in sigmas.items():
        setattr(cfg, k, v)
    outs = [np.zeros_like(raw) for _ in range(3)]
    h, w = raw.shape[:2]
    check(lib().rpd_joint_bilateral(ctx.h, _ptr(raw), _ptr(normal), _ptr(albedo), w, h, C.byref(cfg),
                                    *[_ptr(o) for o in outs], 0))
    return tuple(outs)


def atrous(ctx, raw, normal, albedo, **kw):
    """denoising/atrous_wavelet.cpp on the device, on host images."""
    raw, normal, albedo = _rgb_image(raw), _rgb_image(normal), _rgb_image(albedo)
    cfg = AtrousConfig()
    lib().rpd_atrous_config_defaults(C.byref(cfg))
    for k, v in kw.items():
        setattr(cfg, k, v)
    out = np.zeros_like(raw)
    h, w = raw.shape[:2]
    check(lib().rpd_atrous(ctx.h, _ptr(raw), _ptr(normal), _ptr(albedo), w, h, C.byref(cfg), _ptr(out), 0))
    return out


def statmc_tile_shapes():
    """The StatMC window kernel's tile shapes, [(w, h), ...]."""
    n = C.c_int()
    check(lib().rpd_statmc_tile_shapes(None, 0, C.byref(n)))
    wh = np.zeros(2 * n.value, np.int32)
    check(lib().rpd_statmc_tile_shapes(_ptr(wh), n.value, C.byref(n)))
    return [(int(wh[2 * k]), int(wh[2 * k + 1])) for k in range(n.value)]


def procedural_config(**kw):
    c = ProceduralConfig()
    lib().rpd_procedural_config_defaults(C.byref(c))
    for k, v in kw.items():
        if k == "environment":
            for i in range(3):
                c.environment[i] = v[i]
        else:
            setattr(c, k, v)
    return c


def builtin_scene_arrays(name):
    """(triangles, spheres, materials, lights, env, counts) of a reference scene
    (scenes/scene_registry.h); mesh scenes read resources/models/*.obj from the
    working directory."""
    L = lib()
    n = [C.c_int() for _ in range(4)]
    check(L.rpd_builtin_scene_arrays(name.encode(), None, C.byref(n[0]), None, C.byref(n[1]), None, C.byref(n[2]),
                                     None, C.byref(n[3]), None))
    tris, sph = (Triangle * max(1, n[0].value))(), (Sphere * max(1, n[1].value))()
    mats, lights = (Material * max(1, n[2].value))(), (Light * max(1, n[3].value))()
    env = (C.c_float * 3)()
    check(L.rpd_builtin_scene_arrays(name.encode(), tris, C.byref(n[0]), sph, C.byref(n[1]), mats, C.byref(n[2]),
                                     lights, C.byref(n[3]), env))
    return tris, sph, mats, lights, env, [x.value for x in n]


def obj_triangles(path, material_id=0):
    """Triangles of an OBJ mesh as the engine reads it (core/obj_loader.h)."""
    L = lib()
    n = C.c_int()
    check(L.rpd_obj_load(str(path).encode(), material_id, None, C.byref(n)))
    tris = (Triangle * max(1, n.value))()
    check(L.rpd_obj_load(str(path).encode(), material_id, tris, C.byref(n)))
    return tris, n.value


def procedural_arrays(pc):
    """(triangles, spheres, materials, lights, env) ctypes arrays of the procedural scene."""
    L = lib()
    n = [C.c_int() for _ in range(4)]
    check(L.rpd_procedural_scene_arrays(C.byref(pc), None, C.byref(n[0]), None, C.byref(n[1]), None,
                                        C.byref(n[2]), None, C.byref(n[3]), None))
    tris, sph = (Triangle * max(1, n[0].value))(), (Sphere * max(1, n[1].value))()
    mats, lights = (Material * max(1, n[2].value))(), (Light * max(1, n[3].value))()
    env = (C.c_float * 3)()
    check(L.rpd_procedural_scene_arrays(C.byref(pc), tris, C.byref(n[0]), sph, C.byref(n[1]), mats, C.byref(n[2]),
                                        lights, C.byref(n[3]), env))
    return tris, sph, mats, lights, env, [x.value for x in n]


RPD_BVH_HOST_SAH, RPD_BVH_DEVICE_LBVH = 0, 1


def _scene_query(L, h, origins, dirs, t_max, occlusion, use_bvh):
    o = np.ascontiguousarray(origins, np.float32)
    d = np.ascontiguousarray(dirs, np.float32)
    tm = np.ascontiguousarray(t_max, np.float32)
    oc = np.ascontiguousarray(occlusion, np.int32)
    n = tm.size
    hit, idx, sph = (np.zeros(n, np.int32) for _ in range(3))
    t = np.zeros(n, np.float32)
    check(L.rpd_scene_query(h, n, _ptr(o), _ptr(d), _ptr(tm), _ptr(oc), int(use_bvh), _ptr(hit), _ptr(t), _ptr(idx),
                            _ptr(sph)))
    return hit, t, idx, sph


def _bvh_info(L, h):
    a, b = C.c_int(), C.c_int()
    check(L.rpd_scene_bvh_info(h, C.byref(a), C.byref(b)))
    return a.value, b.value


class HostScene:
    """A scene kept on the host (no device) for rpd_scene_query checks."""

    def __init__(self, pc):
        self.L = lib()
        h = C.c_void_p()
        check(self.L.rpd_scene_procedural(None, C.byref(pc), C.byref(h)))
        self.h = h

    def query(self, origins, dirs, t_max, occlusion, use_bvh):
        return _scene_query(self.L, self.h, origins, dirs, t_max, occlusion, use_bvh)

    def bvh_info(self):
        return _bvh_info(self.L, self.h)

    def close(self):
        if self.h:
            self.L.rpd_scene_destroy(self.h)
            self.h = None


def synth_config(**kw):
    c = SynthConfig()
    lib().rpd_synth_config_defaults(C.byref(c))
    for k, v in kw.items():
        setattr(c, k, v)
    return c


def denoise_batch(ctx, jobs):
    """rpd_denoise_batch over `jobs`; returns the batch's device time in ms."""
    arr = (C.c_void_p * len(jobs))(*[j.h for j in jobs])
    ms = C.c_double()
    check(ctx.L.rpd_denoise_batch(ctx.h, arr, len(jobs), C.byref(ms)))
    return ms.value


def render_finish_batch(ctx, jobs):
    """rpd_render_finish_batch: the denoise tails of jobs submitted with
    submit="passes", as one batch; returns the batch's device time in ms."""
    arr = (C.c_void_p * len(jobs))(*[j.h for j in jobs])
    ms = C.c_double()
    check(ctx.L.rpd_render_finish_batch(ctx.h, arr, len(jobs), C.byref(ms)))
    return ms.value


def job_bytes(cfg):
    b = C.c_uint64()
    check(lib().rpd_job_bytes(C.byref(cfg), C.byref(b)))
    return b.value


class Server:
    """Request loop: continuous admission by slots and HBM footprint (rpd_server_*)."""

    def __init__(self, ctx, max_in_flight=4, memory_budget=0, graph_passes=True, cobatch_tail=False):
        self.ctx = ctx
        self.L = ctx.L
        c = ServerConfig(max_in_flight, int(memory_budget), int(graph_passes), int(cobatch_tail))
        h = C.c_void_p()
        check(self.L.rpd_server_create(ctx.h, C.byref(c), C.byref(h)))
        self.h = h

    def submit(self, cfg, cam, seed, scene="cornell_box"):
        rid = C.c_uint64()
        check(self.L.rpd_server_submit(self.h, self.ctx.scene(scene), C.byref(cfg), C.byref(cam),
                                       seed & 0xFFFFFFFF, C.byref(rid)))
        return rid.value

    def wait(self, rid):
        """Returns (Job, latency ms); the Job owns the finished job."""
        jh = C.c_void_p()
        ms = C.c_double()
        check(self.L.rpd_server_wait(self.h, rid, C.byref(jh), C.byref(ms)))
        return Job.adopt(self.ctx, jh), ms.value

    def stats(self):
        st = ServerStats()
        check(self.L.rpd_server_stats_get(self.h, C.byref(st)))
        return {f: getattr(st, f) for f, _ in ServerStats._fields_}

    def close(self):
        if self.h:
            self.L.rpd_server_destroy(self.h)
            self.h = None


class Job:
    """One render job's device state; stage methods mirror the reference's functions.
    submit=True queues the whole pipeline, submit="passes" the sampling passes only
    (render_finish_batch runs the rest). denoiser_spp > 0 creates a job that
    denoises given records instead (no scene)."""

    def __init__(self, ctx, cfg, cam, seed, scene="cornell_box", submit=False, denoiser_spp=0):
        self.L = ctx.L
        self.cfg = cfg
        self.w, self.h_px = cfg.image_width, cfg.image_height
        h = C.c_void_p()
        if denoiser_spp > 0:
            check(self.L.rpd_job_create_denoiser(ctx.h, C.byref(cfg), int(denoiser_spp), C.byref(h)))
        else:
            fn = (self.L.rpd_job_submit_passes if submit == "passes" else
                  self.L.rpd_job_submit if submit else self.L.rpd_job_create)
            check(fn(ctx.h, ctx.scene(scene), C.byref(cfg), C.byref(cam), seed & 0xFFFFFFFF, C.byref(h)))
        self.h = h
        d = np.zeros(4, np.int32)
        check(self.L.rpd_job_grid_dims(self.h, _ptr(d)))
        self.nodes_x, self.nodes_y, self.step, self.support = (int(v) for v in d)

    @classmethod
    def adopt(cls, ctx, handle):
        """Wraps a job handle the engine handed over (e.g. by rpd_server_wait)."""
        j = cls.__new__(cls)
        j.L = ctx.L
        j.h = handle
        d = np.zeros(4, np.int32)
        check(j.L.rpd_job_grid_dims(j.h, _ptr(d)))
        j.nodes_x, j.nodes_y, j.step, j.support = (int(v) for v in d)
        n = C.c_size_t()
        check(j.L.rpd_job_buffer_size(j.h, BUF["depth"], C.byref(n)))
        j.cfg = None
        j.w, j.h_px = None, None
        j.n_px = n.value // 4
        return j

    def release(self):
        if self.h:
            self.L.rpd_job_release(self.h)
            self.h = None

    def __del__(self):
        try:
            self.release()
        except Exception:
            pass

    def wait(self):
        check(self.L.rpd_job_wait(self.h))

    def poll(self):
        d = C.c_int()
        check(self.L.rpd_job_poll(self.h, C.byref(d)))
        return bool(d.value)

    def run_pass(self, counts=None, statmc=True):
        if counts is None:
            check(self.L.rpd_job_run_pass(self.h, None, int(statmc)))
        else:
            counts = np.ascontiguousarray(counts, np.int32)
            assert counts.size == self.w * self.h_px
            check(self.L.rpd_job_run_pass(self.h, _ptr(counts), int(statmc)))

    def load_records(self, counts, records):
        counts = np.ascontiguousarray(counts, np.int32)
        recs = np.ascontiguousarray(records, SAMPLE_RECORD)
        assert counts.size == self.w * self.h_px
        check(self.L.rpd_job_load_records(self.h, _ptr(counts), _ptr(recs), recs.shape[0]))

    def synthesize_records(self, sc):
        check(self.L.rpd_job_synthesize_records(self.h, C.byref(sc)))

    def denoise_records(self):
        check(self.L.rpd_job_denoise_records(self.h))


    def run_adaptive_pass(self):
        check(self.L.rpd_job_run_adaptive_pass(self.h))

    def recompute_sensitivity(self):
        check(self.L.rpd_job_recompute_sensitivity(self.h))

    def statmc(self):
        check(self.L.rpd_job_statmc_denoise(self.h))

    def variance(self):
        check(self.L.rpd_job_variance_denoise(self.h))

    def adaptive_counts(self):
        out = np.zeros(self.w * self.h_px, np.int32)
        check(self.L.rpd_job_adaptive_counts(self.h, _ptr(out)))
        return out

    def counts_from_importance(self, importance):
        imp = np.ascontiguousarray(importance, np.float32)
        assert imp.size == self.w * self.h_px
        out = np.zeros(self.w * self.h_px, np.int32)
        check(self.L.rpd_job_adaptive_counts_from_importance(self.h, _ptr(imp), _ptr(out)))
        return out

    def finalize(self):
        check(self.L.rpd_job_finalize(self.h))

    def stage_ms(self):
        out = np.zeros(11, np.float64)
        check(self.L.rpd_job_stage_ms(self.h, _ptr(out), 11))
        return out

    def counters(self):
        """(ray queries, path samples, triangles, spheres) traced by this job so far."""
        out = np.zeros(4, np.uint64)
        check(self.L.rpd_job_counters(self.h, _ptr(out), 4))
        return dict(queries=int(out[0]), samples=int(out[1]), triangles=int(out[2]), spheres=int(out[3]))

    def nbytes(self, name):
        n = C.c_size_t()
        check(self.L.rpd_job_buffer_size(self.h, BUF[name], C.byref(n)))
        return n.value

    def metrics(self, name, ref):
        """Quality metrics of this job's `name` buffer (radiance, raw_radiance or
        denoised) against the host image `ref` ((H, W, 3) or (N, 3) float32)."""
        ref = np.ascontiguousarray(ref, np.float32)
        m = ImageMetrics()
        check(self.L.rpd_job_image_metrics(self.h, BUF[name], _ptr(ref), 0, C.byref(m)))
        return m.as_dict()

    def get(self, name):
        nb = self.nbytes(name)
        if name == "grid":
            out = np.zeros(nb // 8, np.float64)
        elif name in _INT:
            out = np.zeros(nb // 4, np.int32)
        else:
            out = np.zeros(max(1, nb // 4), np.float32)
        check(self.L.rpd_job_read(self.h, BUF[name], _ptr(out), nb))
        if name == "grid":
            return out.reshape(-1, 440)
        if name == "tiles":
            return out.reshape(18, -1)
        if name in _RGB:
            return out.reshape(-1, 3)
        if name in _WIDTH:
            return out[: nb // 4].reshape(-1, _WIDTH[name])
        return out

    def set(self, name, arr):
        if name == "grid":
            arr = np.ascontiguousarray(arr, np.float64)
        elif name in _INT:
            arr = np.ascontiguousarray(arr, np.int32)
        else:
            arr = np.ascontiguousarray(arr, np.float32)
        check(self.L.rpd_job_write(self.h, BUF[name], _ptr(arr), arr.nbytes))
