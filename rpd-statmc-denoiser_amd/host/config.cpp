#include "config.h"

#include <algorithm>
#include <array>
#include <cmath>
#include <stdexcept>

namespace host {

namespace {
const std::array<const char*, 33> kKnownKeys = {
    "image_width", "image_height", "samples_per_pixel", "fov", "focus_distance", "aperture",
    "camera_position", "camera_direction", "world_up", "scene", "output", "statmc_enabled",
    "rpf_tile_size", "color_window_radius", "color_normal_threshold", "color_depth_threshold",
    "color_compat_alpha", "color_sigma_max", "var_window_radius", "var_normal_threshold",
    "var_depth_threshold", "var_compat_sigma", "var_shrinkage_k", "var_iterations",
    "adaptive_base_samples", "adaptive_spp", "adaptive_sigma_max", "adaptive_passes",
    "adaptive_importance_smoothing_radius", "tonemap", "rpf_shrinkage_scale",
    "rpf_confidence_samples", "debug_statmc_outputs"};

bool finite3(const float* v) { return std::isfinite(v[0]) && std::isfinite(v[1]) && std::isfinite(v[2]); }
float sq3(const float* v) { return v[0] * v[0] + (v[1] * v[1] + v[2] * v[2]); }
}  // namespace

// core/config.h:51-201
Settings parse_settings(const FlatYaml& y) {
    for (const auto& key : y.keys_in_order()) {
        if (std::find_if(kKnownKeys.begin(), kKnownKeys.end(), [&](const char* k) { return key == k; }) == kKnownKeys.end())
            throw std::runtime_error("Unknown render config key: " + key);
    }
    Settings s;
    rpd_render_config_defaults(&s.render);
    rpd_render_config& r = s.render;
    r.image_width = y.as_int("image_width");
    r.image_height = y.as_int("image_height");
    r.samples_per_pixel = y.as_int("samples_per_pixel");
    s.fov = y.as_float("fov");
    const std::vector<float> pos = y.as_floats("camera_position");
    const std::vector<float> dir = y.as_floats("camera_direction");
    const std::vector<float> up = y.has("world_up") ? y.as_floats("world_up") : std::vector<float>{0.0f, 1.0f, 0.0f};
    if (pos.size() != 3 || dir.size() != 3) throw std::runtime_error("Camera vectors must have three components");
    if (up.size() != 3) throw std::runtime_error("world_up must have three components");
    for (int k = 0; k < 3; ++k) { s.camera_position[k] = pos[k]; s.camera_direction[k] = dir[k]; s.camera_up[k] = up[k]; }
    // focus_distance defaults to the length of camera_direction (config.h:90)
    s.focus_distance = y.has("focus_distance") ? y.as_float("focus_distance") : std::sqrt(sq3(s.camera_direction));
    s.aperture = y.has("aperture") ? y.as_float("aperture") : 0.0f;
    if (y.has("scene")) s.scene_name = y.as_string("scene");
    if (y.has("output")) s.output_dir = y.as_string("output");
    if (y.has("statmc_enabled")) r.use_statmc = y.as_bool("statmc_enabled") ? 1 : 0;
    auto geti = [&](const char* k, int& dst) { if (y.has(k)) dst = y.as_int(k); };
    auto getf = [&](const char* k, float& dst) { if (y.has(k)) dst = y.as_float(k); };
    geti("rpf_tile_size", r.rpf_tile_size);
    geti("color_window_radius", r.color_window_radius);
    getf("color_normal_threshold", r.color_normal_threshold);
    getf("color_depth_threshold", r.color_depth_threshold);
    getf("color_compat_alpha", r.color_compat_alpha);
    getf("color_sigma_max", r.color_sigma_max);
    geti("var_window_radius", r.var_window_radius);
    getf("var_normal_threshold", r.var_normal_threshold);
    getf("var_depth_threshold", r.var_depth_threshold);
    getf("var_compat_sigma", r.var_compat_sigma);
    getf("var_shrinkage_k", r.var_shrinkage_k);
    geti("var_iterations", r.var_iterations);
    geti("adaptive_base_samples", r.adaptive_base_samples);
    geti("adaptive_spp", r.adaptive_spp);
    getf("adaptive_sigma_max", r.adaptive_sigma_max);
    geti("adaptive_passes", r.adaptive_passes);
    geti("adaptive_importance_smoothing_radius", r.adaptive_importance_smoothing_radius);
    getf("rpf_shrinkage_scale", r.rpf_shrinkage_scale);
    getf("rpf_confidence_samples", r.rpf_confidence_samples);
    if (y.has("debug_statmc_outputs")) s.debug_statmc_outputs = y.as_bool("debug_statmc_outputs");

    auto bad = [](const char* m) { throw std::runtime_error(m); };
    if (r.image_width <= 0 || r.image_height <= 0) bad("image_width and image_height must be positive");
    if (r.samples_per_pixel <= 0) bad("samples_per_pixel must be positive");
    if (!std::isfinite(s.fov) || s.fov <= 0.0f || s.fov >= 180.0f) bad("fov must be finite and in (0,180)");
    if (!finite3(s.camera_position) || !finite3(s.camera_direction) || !finite3(s.camera_up) ||
        sq3(s.camera_direction) <= 1e-8f || sq3(s.camera_up) <= 1e-8f)
        bad("Camera vectors must be finite and direction/up must be nonzero");
    if (!std::isfinite(s.focus_distance) || s.focus_distance <= 0.0f) bad("focus_distance must be finite and positive");
    if (!std::isfinite(s.aperture) || s.aperture < 0.0f) bad("aperture must be finite and non-negative");
    if (r.rpf_tile_size <= 0) bad("rpf_tile_size must be positive");
    if (r.color_window_radius <= 0) bad("color_window_radius must be positive");
    if (!std::isfinite(r.color_normal_threshold) || r.color_normal_threshold <= 0.0f || r.color_normal_threshold > 1.0f)
        bad("color_normal_threshold must be finite and in (0,1]");
    if (!std::isfinite(r.color_depth_threshold) || r.color_depth_threshold < 0.0f)
        bad("color_depth_threshold must be finite and non-negative");
    if (!std::isfinite(r.color_compat_alpha) || r.color_compat_alpha <= 0.0f || r.color_compat_alpha >= 1.0f)
        bad("color_compat_alpha must be finite and in (0,1)");
    if (!std::isfinite(r.color_sigma_max) || r.color_sigma_max <= 0.0f) bad("color_sigma_max must be finite and positive");
    if (r.var_window_radius <= 0) bad("var_window_radius must be positive");
    if (!std::isfinite(r.var_normal_threshold) || r.var_normal_threshold <= 0.0f || r.var_normal_threshold > 1.0f)
        bad("var_normal_threshold must be finite and in (0,1]");
    if (!std::isfinite(r.var_depth_threshold) || r.var_depth_threshold < 0.0f)
        bad("var_depth_threshold must be finite and non-negative");
    if (!std::isfinite(r.var_compat_sigma) || r.var_compat_sigma <= 0.0f) bad("var_compat_sigma must be finite and positive");
    if (!std::isfinite(r.var_shrinkage_k) || r.var_shrinkage_k <= 0.0f) bad("var_shrinkage_k must be finite and positive");
    if (r.adaptive_base_samples < 0) bad("adaptive_base_samples must be non-negative");
    if (r.adaptive_spp < 0) bad("adaptive_spp must be non-negative");
    if (r.adaptive_base_samples > r.adaptive_spp) bad("adaptive_base_samples cannot exceed adaptive_spp");
    if (!std::isfinite(r.adaptive_sigma_max) || r.adaptive_sigma_max <= 0.0f) bad("adaptive_sigma_max must be finite and positive");
    if (r.adaptive_passes < 0) bad("adaptive_passes must be non-negative");
    if (r.adaptive_importance_smoothing_radius < 0) bad("adaptive_importance_smoothing_radius must be non-negative");
    if (r.var_iterations <= 0) bad("var_iterations must be positive");
    if (!std::isfinite(r.rpf_shrinkage_scale) || r.rpf_shrinkage_scale < 0.0f) bad("rpf_shrinkage_scale must be finite and non-negative");
    if (!std::isfinite(r.rpf_confidence_samples) || r.rpf_confidence_samples <= 0.0f)
        bad("rpf_confidence_samples must be finite and positive");
    if (y.has("tonemap")) s.tonemap = y.as_string("tonemap");
    return s;
}

}  // namespace host
