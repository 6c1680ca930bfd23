// ORACLE — test infrastructure only (see orc_common.h).
// Per-job state of the CPU restatement. It mirrors the reference's process
// globals (path_tracer/main.cpp:38-79) and FrameBuffers (core/buffers.h:12-168).
#pragma once
#include <array>
#include <random>
#include <vector>
#include <string>
#include "orc_scene.h"

namespace orc {

// core/config.h:15-49 (defaults identical)
struct Config {
    int width = 0, height = 0, spp = 1;
    float fov = 45.0f, focus_distance = 1.0f, aperture = 0.0f;
    F3 cam_pos{0, 0, 0}, cam_dir{0, 0, 1}, cam_up{0, 1, 0};
    int use_statmc = 0;
    int rpf_tile_size = 8;
    int color_window_radius = 7;
    float color_normal_threshold = 0.5f, color_depth_threshold = 0.25f;
    float color_compat_alpha = 0.05f, color_sigma_max = 6.0f;
    int var_window_radius = 1;
    float var_normal_threshold = 0.5f, var_depth_threshold = 0.25f;
    float var_compat_sigma = 30.0f, var_shrinkage_k = 1e-3f;
    int var_iterations = 2;
    int adaptive_base_samples = 0, adaptive_spp = 0;
    float adaptive_sigma_max = 3.0f;
    int adaptive_passes = 1, adaptive_importance_smoothing_radius = 1;
    float rpf_shrinkage_scale = 1.0f, rpf_confidence_samples = 128.0f;
    // extension of the engine (not in the reference): shading points per path, 0 = unbounded
    int max_depth = 0;
};

// core/camera.h:9-88
struct Camera {
    F3 position, forward{0, 0, 1}, right{1, 0, 0}, up{0, 1, 0};
    float fov = 45.0f, aspect = 1.0f, focus = 1.0f, lens_radius = 0.0f;
    F3 llc, horizontal, vertical;
    void init(F3 eye, F3 dir, float fov_deg, float aspect_ratio, float focus_dist, float aperture, F3 world_up);
    Ray generate(float u, float v, F2 lens) const;
};

// core/sampler.h:6-44: libstdc++ mt19937 + uniform_real_distribution<float>.
// Pairs are (second draw, first draw): GCC evaluates the two next1d() calls of
// Vec2f(next1d(), next1d()) right to left.
struct Sampler {
    std::mt19937 rng;
    std::uniform_real_distribution<float> dist{0.0f, 1.0f};
    void seed(uint32_t s) { rng.seed(s); }
    float next1d() { return dist(rng); }
    F2 next2d() { const float first = next1d(); const float second = next1d(); return {second, first}; }
};
F2 sample_disk(F2 u);
F3 sample_sphere(F2 u);

// core/stats.h:6-41
struct PixelStats {
    float mean = 0.0f, square_diff = 0.0f;
    int n = 0;
    F3 color_mean, sqrt_color_mean, sqrt_color_square_diff;
    float variance() const { return (n > 1) ? square_diff / float(n - 1) : 0.0f; }
    F3 sqrt_color_variance() const { return (n <= 1) ? F3{0, 0, 0} : sqrt_color_square_diff / float(n - 1); }
};
void accumulate_sample(PixelStats& s, F3 c);

// core/stats.h:93-159
template <int N>
struct Binned {
    std::array<double, N> weight{}, sum_y{};
    double w_sum = 0, w2_sum = 0, y_sum = 0, y2_sum = 0;
    void add(int bin, float y, float w) {
        if (w <= ORC_EPS_SMALL) return;
        const int b = std::clamp(bin, 0, N - 1);
        const double wd = double(w), yd = double(y);
        weight[b] += wd; sum_y[b] += wd * yd;
        w_sum += wd; w2_sum += wd * wd; y_sum += wd * yd; y2_sum += wd * yd * yd;
    }
    double mass() const { return w2_sum > 0.0 ? w_sum * w_sum / w2_sum : 0.0; }
    float sensitivity() const;
    void merge(const Binned& o) {
        for (int b = 0; b < N; ++b) { weight[b] += o.weight[b]; sum_y[b] += o.sum_y[b]; }
        w_sum += o.w_sum; w2_sum += o.w2_sum; y_sum += o.y_sum; y2_sum += o.y2_sum;
    }
};
// core/stats.h:161-182
struct Dep2D {
    Binned<16> b;
    void add(float u, float v, float y, float w) {
        const int x = std::clamp(int(u * 4), 0, 3), z = std::clamp(int(v * 4), 0, 3);
        b.add(z * 4 + x, y, w);
    }
};
// core/stats.h:184-303; the 13 accumulators in declaration order.
enum TileField {
    TF_PIXEL = 0, TF_BRDF, TF_SCREEN_BRDF, TF_LENS, TF_SCREEN_LENS, TF_LIGHT_UV, TF_SCREEN_LIGHT_UV,
    TF_ENV, TF_SCREEN_ENV, TF_LIGHT_SELECT, TF_SCREEN_LIGHT_SELECT, TF_RR, TF_SCREEN_RR, TF_COUNT
};
struct Tile {
    Dep2D pixel, brdf, screen_brdf, lens, screen_lens, light_uv, screen_light_uv, env, screen_env;
    Binned<16> light_select;
    Dep2D screen_light_select;
    Binned<2> rr;
    Dep2D screen_rr;
    void merge(const Tile& o);
    // sensitivity and confidence of each accumulator, indexed by TileField
    void split(float n0, float sens[TF_COUNT], float conf[TF_COUNT]) const;
};
// core/stats.h:305-329
struct Grid {
    int support = 1, step = 1, nx = 0, ny = 0;
    std::vector<Tile> nodes;
    void init(int w, int h, int size);
};

// core/stats.h:331-399
struct RpfSample {
    F2 pixel_u, lens_u, brdf_u, light_u, env_u;
    int light_index = -1;
    bool rr_survived = false;
    bool pixel_valid = false, lens_valid = false, light_select_valid = false, brdf_valid = false,
         light_valid = false, env_valid = false, rr_valid = false;
    bool any_shading() const { return light_select_valid || brdf_valid || light_valid || env_valid || rr_valid; }
};

// One traced sample as the tests see it (per-sample record).
struct SampleRecord {
    int pixel = 0;
    F3 radiance;
    int finite = 0;
    int hit = 0;
    F3 normal, albedo, position;
    float depth = 0.0f;
    RpfSample rpf;
};

// Node-level sensitivity tiles, core/buffers.h:18-66. Order of fields:
enum SensField {
    SF_ALL = 0, SF_ALL_CONF, SF_PIXEL, SF_PIXEL_CONF, SF_BRDF, SF_BRDF_CONF, SF_LENS, SF_LENS_CONF,
    SF_LIGHT, SF_LIGHT_CONF, SF_LIGHT_UV, SF_LIGHT_UV_CONF, SF_LIGHT_SELECT, SF_LIGHT_SELECT_CONF,
    SF_ENV, SF_ENV_CONF, SF_RR, SF_RR_CONF, SF_COUNT
};

struct Job {
    Config cfg;
    Scene scene;
    Camera cam;
    uint32_t seed = 0;
    int nthreads = 1;
    bool skip_dead_work = false;   // skip the reference's overwritten in-loop calls
    bool keep_records = false;
    bool verbose = false;
    int w = 0, h = 0;
    // FrameBuffers
    std::vector<F3> radiance, raw_radiance, albedo, normal, world_pos, denoised;
    std::vector<float> depth, uncertainty, alpha, variance_denoised, var_total, var_eff, var_ratio;
    std::vector<float> sens_confidence, sens_gradient, adaptive_importance;
    std::vector<int> hit_count, sample_count;
    std::vector<PixelStats> stats;
    Grid grid;
    int tiles_support = 1, tiles_step = 1, tiles_nx = 0, tiles_ny = 0;
    std::vector<float> tiles[SF_COUNT];
    // derived caches (main.cpp:45-154)
    bool feat_valid = false;
    std::vector<F3> feat_normal, feat_albedo;
    std::vector<float> feat_depth;
    bool sens_valid = false;
    std::vector<float> cache_comp[6];  // brdf, lens, light_uv, light_select, environment, rr
    std::vector<float> cache_grad;
    // last pass
    std::vector<int> extra_counts;
    std::vector<SampleRecord> records;
    double stage_ms[8] = {0};  // render, adaptive render, sensitivity, statmc, variance, allocation, total, -
};

void job_init_buffers(Job& j);
void render_pass(Job& j, const int* counts, bool statmc, bool seed_with_n);
// Folds given per-sample records (pixel-major, counts[p] per pixel) exactly as
// a StatMC pass folds the samples it traces (main.cpp:707-717).
void accumulate_records(Job& j, const int* counts, const SampleRecord* recs);
void recompute_sensitivity(Job& j);
void statmc_denoise(Job& j);
void variance_denoise(Job& j);
void compute_adaptive_counts(Job& j, std::vector<int>& counts);
void counts_from_importance(Job& j, const std::vector<float>& imp, double sum, std::vector<int>& counts);
void finalize(Job& j);
void render(Job& j);
const std::vector<F3>& feature_normals(Job& j);
void ensure_features(Job& j);
void ensure_sensitivity_cache(Job& j);
float expand_lookup(const Job& j, const std::vector<float>& tiles, int x, int y);
float lookup_t_critical(float df, float alpha);

}  // namespace orc
