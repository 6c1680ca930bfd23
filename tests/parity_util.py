"""Shared helpers for GPU-vs-oracle parity tests."""
import numpy as np

import oracle_bind as ob
import rpd_engine as eng


def engine_config(rc: ob.RenderConfig):
    """rpd_render_config from the oracle's RenderConfig dataclass."""
    return eng.default_config(
        image_width=rc.width, image_height=rc.height, samples_per_pixel=rc.spp, use_statmc=rc.use_statmc,
        rpf_tile_size=rc.rpf_tile_size, color_window_radius=rc.color_window_radius,
        color_normal_threshold=rc.color_normal_threshold, color_depth_threshold=rc.color_depth_threshold,
        color_compat_alpha=rc.color_compat_alpha, color_sigma_max=rc.color_sigma_max,
        var_window_radius=rc.var_window_radius, var_normal_threshold=rc.var_normal_threshold,
        var_depth_threshold=rc.var_depth_threshold, var_compat_sigma=rc.var_compat_sigma,
        var_shrinkage_k=rc.var_shrinkage_k, var_iterations=rc.var_iterations,
        adaptive_base_samples=rc.adaptive_base_samples, adaptive_spp=rc.adaptive_spp,
        adaptive_sigma_max=rc.adaptive_sigma_max, adaptive_passes=rc.adaptive_passes,
        adaptive_importance_smoothing_radius=rc.adaptive_importance_smoothing_radius,
        rpf_shrinkage_scale=rc.rpf_shrinkage_scale, rpf_confidence_samples=rc.rpf_confidence_samples,
        max_depth=rc.max_depth)


def engine_camera(rc: ob.RenderConfig):
    return eng.make_camera(rc.cam_pos, rc.cam_dir, rc.cam_up, rc.fov, rc.focus_distance, rc.aperture)


def rel_err(a, b, floor=1e-6):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return np.abs(a - b) / np.maximum(np.maximum(np.abs(a), np.abs(b)), floor)
