// path_tracer: the reference renderer's command line (path_tracer/main.cpp:1566-1922)
// on top of the MI355X engine's C ABI. Same flags, YAML keys, validation
// messages, output directory rules, output files and stage log lines.
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <ctime>
#include <filesystem>
#include <limits>
#include <map>
#include <optional>
#include <random>
#include <string>
#include <vector>

#include "../../include/rpd_engine.h"
#include "config.h"
#include "flat_yaml.h"
#include "image_io.h"

namespace fs = std::filesystem;

namespace {

void log_line(const char* level, const char* fmt, ...) {
    char msg[4096];
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(msg, sizeof(msg), fmt, ap);
    va_end(ap);
    const auto now = std::chrono::system_clock::now();
    const std::time_t t = std::chrono::system_clock::to_time_t(now);
    const int ms = int(std::chrono::duration_cast<std::chrono::milliseconds>(now.time_since_epoch()).count() % 1000);
    std::tm tmv;
    localtime_r(&t, &tmv);
    char ts[64];
    std::strftime(ts, sizeof(ts), "%Y-%m-%d %H:%M:%S", &tmv);
    std::FILE* out = std::strcmp(level, "error") == 0 ? stderr : stdout;
    std::fprintf(out, "[%s.%03d] [%s] %s\n", ts, ms, level, msg);
    std::fflush(out);
}
#define INFO(...) log_line("info", __VA_ARGS__)
#define ERROR(...) log_line("error", __VA_ARGS__)

// ---- command line (CLI11 semantics of main.cpp:1597-1625) ------------------------------
enum CliExit { CLI_OK = 0, CLI_CONVERSION = 104, CLI_VALIDATION = 105, CLI_REQUIRED = 106, CLI_EXTRAS = 109 };

struct Cli {
    std::string config_file, scene_name, output_dir, tonemap;
    bool statmc_on = false, statmc_off = false;
    std::map<std::string, std::string> values;  // option name -> raw value
    std::optional<uint32_t> seed;
};

const char* kValueOptions[] = {"--spp", "--rpf-tile-size", "--color-window-radius", "--color-normal-thresh",
                               "--color-depth-thresh", "--color-compat-alpha", "--color-sigma-max",
                               "--var-window-radius", "--var-normal-thresh", "--var-depth-thresh",
                               "--var-compat-sigma", "--var-shrinkage-k", "--var-iterations",
                               "--adaptive-base-samples", "--adaptive-spp", "--adaptive-sigma-max",
                               "--adaptive-passes", "--adaptive-importance-radius", "--rpf-shrinkage-scale",
                               "--rpf-confidence-samples"};
const char* kPositive[] = {"--spp", "--rpf-tile-size", "--color-window-radius", "--var-window-radius",
                           "--var-iterations"};

void usage() {
    std::printf(
        "Monte Carlo path tracer\nUsage: path_tracer [OPTIONS]\n\nOptions:\n"
        "  -h,--help                   Print this help message and exit\n"
        "  -c,--config TEXT REQUIRED   Path to config YAML file\n"
        "  -s,--scene TEXT             Name of the built-in scene to render (overrides config)\n"
        "  -o,--output TEXT            Output directory (absolute or relative). Image saved as <dir>/<name>.png\n"
        "  --spp INT                   Override samples per pixel (>0)\n"
        "  --statmc / --no-statmc      Enable / disable StatMC/RPF rendering\n"
        "  --rpf-tile-size INT         Sensitivity support size for the overlapping RPF grid\n"
        "  --color-window-radius INT, --color-normal-thresh, --color-depth-thresh, --color-compat-alpha,\n"
        "  --color-sigma-max, --var-window-radius INT, --var-normal-thresh, --var-depth-thresh,\n"
        "  --var-compat-sigma, --var-shrinkage-k, --var-iterations INT, --adaptive-base-samples INT,\n"
        "  --adaptive-spp INT, --adaptive-sigma-max, --adaptive-passes INT, --adaptive-importance-radius INT,\n"
        "  --rpf-shrinkage-scale, --rpf-confidence-samples\n"
        "  --tonemap TEXT              Tonemapping preset: aces | agx | agx-golden | agx-punchy\n"
        "  --seed UINT                 Optional RNG seed for the sampler\n");
}

int parse_cli(int argc, char** argv, Cli& c, bool& help) {
    help = false;
    bool have_config = false;
    for (int i = 1; i < argc; ++i) {
        std::string a = argv[i], val;
        bool has_eq = false;
        if (a.rfind("--", 0) == 0) {
            const size_t eq = a.find('=');
            if (eq != std::string::npos) { val = a.substr(eq + 1); a = a.substr(0, eq); has_eq = true; }
        }
        auto take = [&](std::string& dst) -> int {
            if (has_eq) { dst = val; return CLI_OK; }
            if (i + 1 >= argc) { std::fprintf(stderr, "%s: 1 required TEXT missing\n", a.c_str()); return CLI_CONVERSION; }
            dst = argv[++i];
            return CLI_OK;
        };
        int rc = CLI_OK;
        if (a == "-h" || a == "--help") { help = true; return CLI_OK; }
        else if (a == "-c" || a == "--config") { rc = take(c.config_file); have_config = true; }
        else if (a == "-s" || a == "--scene") rc = take(c.scene_name);
        else if (a == "-o" || a == "--output") rc = take(c.output_dir);
        else if (a == "--tonemap") rc = take(c.tonemap);
        else if (a == "--statmc") c.statmc_on = true;
        else if (a == "--no-statmc") c.statmc_off = true;
        else if (a == "--seed") {
            std::string s;
            rc = take(s);
            if (!rc) {
                char* end = nullptr;
                const unsigned long long v = std::strtoull(s.c_str(), &end, 10);
                if (s.empty() || *end || s[0] == '-' || v > 0xffffffffull) {
                    std::fprintf(stderr, "Could not convert: --seed = %s\n", s.c_str());
                    return CLI_CONVERSION;
                }
                c.seed = uint32_t(v);
            }
        } else {
            bool known = false;
            for (const char* o : kValueOptions)
                if (a == o) { known = true; break; }
            if (!known) {
                std::fprintf(stderr, "The following arguments were not expected: %s\nRun with --help for more information.\n", argv[i]);
                return CLI_EXTRAS;
            }
            std::string s;
            rc = take(s);
            if (!rc) c.values[a] = s;
        }
        if (rc) return rc;
    }
    if (!have_config) {
        std::fprintf(stderr, "--config is required\nRun with --help for more information.\n");
        return CLI_REQUIRED;
    }
    // numeric conversions and PositiveNumber checks
    for (const auto& kv : c.values) {
        bool is_int = false;
        for (const char* o : {"--spp", "--rpf-tile-size", "--color-window-radius", "--var-window-radius",
                              "--var-iterations", "--adaptive-base-samples", "--adaptive-spp", "--adaptive-passes",
                              "--adaptive-importance-radius"})
            if (kv.first == o) is_int = true;
        char* end = nullptr;
        const double v = is_int ? double(std::strtol(kv.second.c_str(), &end, 10)) : std::strtod(kv.second.c_str(), &end);
        if (kv.second.empty() || *end) {
            std::fprintf(stderr, "Could not convert: %s = %s\n", kv.first.c_str(), kv.second.c_str());
            return CLI_CONVERSION;
        }
        for (const char* o : kPositive)
            if (kv.first == o && !(v > 0)) {
                std::fprintf(stderr, "%s: Value %s not in range [0 - inf) or not positive\n", kv.first.c_str(), kv.second.c_str());
                return CLI_VALIDATION;
            }
    }
    return CLI_OK;
}

bool has(const Cli& c, const char* k) { return c.values.count(k) != 0; }
int geti(const Cli& c, const char* k) { return int(std::strtol(c.values.at(k).c_str(), nullptr, 10)); }
float getf(const Cli& c, const char* k) { return std::strtof(c.values.at(k).c_str(), nullptr); }

#define CHECK(rcall)                                                                \
    do {                                                                            \
        const int rc_ = (rcall);                                                    \
        if (rc_ != 0) {                                                             \
            ERROR("%s failed: %s", #rcall, rpd_last_error());                       \
            return 1;                                                               \
        }                                                                           \
    } while (0)

}  // namespace

int main(int argc, char** argv) {
    Cli cli;
    bool help = false;
    const int crc = parse_cli(argc, argv, cli, help);
    if (help) { usage(); return 0; }
    if (crc) return crc;

    host::FlatYaml yaml;
    try {
        yaml = host::FlatYaml::load_file(cli.config_file);
    } catch (const std::exception& e) {
        ERROR("%s", e.what());
        return 1;
    }
    host::Settings st;
    try {
        st = host::parse_settings(yaml);
    } catch (const std::exception& e) {
        ERROR("Failed to parse render config: %s", e.what());
        return 1;
    }
    if (cli.statmc_on && cli.statmc_off) { ERROR("Cannot specify both --statmc and --no-statmc"); return 1; }
    const std::string scene = !cli.scene_name.empty() ? cli.scene_name : st.scene_name.value_or("");
    if (scene.empty()) { ERROR("No scene specified. Provide --scene or set 'scene' in the config."); return 1; }

    // output directory resolution, main.cpp:1654-1670
    fs::path out_dir;
    if (!cli.output_dir.empty()) {
        out_dir = fs::path(cli.output_dir);
        if (st.output_dir && cli.output_dir != *st.output_dir)
            INFO("CLI output '%s' overrides config output '%s'", cli.output_dir.c_str(), st.output_dir->c_str());
    } else if (st.output_dir) {
        out_dir = fs::path(*st.output_dir);
        if (!out_dir.is_absolute()) out_dir = fs::path("output") / out_dir;
    } else {
        ERROR("No output directory specified. Provide --output or set 'output' in the config.");
        return 1;
    }

    // CLI over YAML, with the reference's re-validation (main.cpp:1672-1836)
    rpd_render_config& r = st.render;
    if (has(cli, "--spp")) {
        const int v = geti(cli, "--spp");
        if (v <= 0) { ERROR("Samples per pixel must be positive"); return 1; }
        INFO("CLI samples-per-pixel %d overrides config %d", v, r.samples_per_pixel);
        r.samples_per_pixel = v;
    }
    if (cli.statmc_on) r.use_statmc = 1;
    if (cli.statmc_off) r.use_statmc = 0;
    if (has(cli, "--rpf-tile-size")) r.rpf_tile_size = geti(cli, "--rpf-tile-size");
    if (r.rpf_tile_size <= 0) { ERROR("RPF tile size must be positive"); return 1; }
    if (has(cli, "--color-window-radius")) r.color_window_radius = geti(cli, "--color-window-radius");
    if (r.color_window_radius <= 0) { ERROR("color_window_radius must be positive"); return 1; }
    if (has(cli, "--color-normal-thresh")) r.color_normal_threshold = getf(cli, "--color-normal-thresh");
    if (r.color_normal_threshold <= 0.0f || r.color_normal_threshold > 1.0f) { ERROR("color_normal_thresh must be in (0,1]"); return 1; }
    if (has(cli, "--color-depth-thresh")) r.color_depth_threshold = getf(cli, "--color-depth-thresh");
    if (r.color_depth_threshold < 0.0f) { ERROR("color_depth_thresh must be non-negative"); return 1; }
    if (has(cli, "--color-compat-alpha")) r.color_compat_alpha = getf(cli, "--color-compat-alpha");
    if (r.color_compat_alpha <= 0.0f || r.color_compat_alpha >= 1.0f) { ERROR("color_compat_alpha must be in (0,1)"); return 1; }
    if (has(cli, "--color-sigma-max")) r.color_sigma_max = getf(cli, "--color-sigma-max");
    if (r.color_sigma_max <= 0.0f) { ERROR("color_sigma_max must be positive"); return 1; }
    if (has(cli, "--var-window-radius")) r.var_window_radius = geti(cli, "--var-window-radius");
    if (r.var_window_radius <= 0) { ERROR("var_window_radius must be positive"); return 1; }
    if (has(cli, "--var-normal-thresh")) r.var_normal_threshold = getf(cli, "--var-normal-thresh");
    if (r.var_normal_threshold <= 0.0f || r.var_normal_threshold > 1.0f) { ERROR("var_normal_thresh must be in (0,1]"); return 1; }
    if (has(cli, "--var-depth-thresh")) r.var_depth_threshold = getf(cli, "--var-depth-thresh");
    if (r.var_depth_threshold < 0.0f) { ERROR("var_depth_thresh must be non-negative"); return 1; }
    if (has(cli, "--var-compat-sigma")) r.var_compat_sigma = getf(cli, "--var-compat-sigma");
    if (r.var_compat_sigma <= 0.0f) { ERROR("var_compat_sigma must be positive"); return 1; }
    if (has(cli, "--var-shrinkage-k")) r.var_shrinkage_k = getf(cli, "--var-shrinkage-k");
    if (r.var_shrinkage_k <= 0.0f) { ERROR("var_shrinkage_k must be positive"); return 1; }
    if (has(cli, "--var-iterations")) r.var_iterations = geti(cli, "--var-iterations");
    if (r.var_iterations <= 0) { ERROR("var_iterations must be positive"); return 1; }
    if (has(cli, "--adaptive-base-samples")) r.adaptive_base_samples = geti(cli, "--adaptive-base-samples");
    if (has(cli, "--adaptive-spp")) r.adaptive_spp = geti(cli, "--adaptive-spp");
    if (has(cli, "--adaptive-sigma-max")) r.adaptive_sigma_max = getf(cli, "--adaptive-sigma-max");
    if (has(cli, "--adaptive-passes")) r.adaptive_passes = geti(cli, "--adaptive-passes");
    if (has(cli, "--adaptive-importance-radius")) r.adaptive_importance_smoothing_radius = geti(cli, "--adaptive-importance-radius");
    if (r.adaptive_base_samples < 0 || r.adaptive_spp < 0) { ERROR("adaptive_base_samples and adaptive_spp must be non-negative"); return 1; }
    if (r.adaptive_base_samples > r.adaptive_spp) { ERROR("adaptive_base_samples cannot exceed adaptive_spp"); return 1; }
    if (r.adaptive_sigma_max <= 0.0f) { ERROR("adaptive_sigma_max must be positive"); return 1; }
    if (r.adaptive_passes < 0) { ERROR("adaptive_passes must be non-negative"); return 1; }
    if (r.adaptive_importance_smoothing_radius < 0) { ERROR("adaptive_importance_smoothing_radius must be non-negative"); return 1; }
    if (has(cli, "--rpf-shrinkage-scale")) r.rpf_shrinkage_scale = getf(cli, "--rpf-shrinkage-scale");
    if (!std::isfinite(r.rpf_shrinkage_scale) || r.rpf_shrinkage_scale < 0.0f) { ERROR("rpf_shrinkage_scale must be finite and non-negative"); return 1; }
    if (has(cli, "--rpf-confidence-samples")) r.rpf_confidence_samples = getf(cli, "--rpf-confidence-samples");
    if (!std::isfinite(r.rpf_confidence_samples) || r.rpf_confidence_samples <= 0.0f) { ERROR("rpf_confidence_samples must be finite and positive"); return 1; }
    std::string tonemap_name = st.tonemap;
    if (!cli.tonemap.empty()) {
        if (st.tonemap != cli.tonemap) INFO("CLI tonemap '%s' overrides config tonemap '%s'", cli.tonemap.c_str(), st.tonemap.c_str());
        tonemap_name = cli.tonemap;
    }
    host::ToneMapping tm;
    if (!host::tonemap_from_string(tonemap_name, tm)) {
        ERROR("Unknown tonemap preset '%s'. Allowed: aces, agx, agx-golden, agx-punchy", tonemap_name.c_str());
        return 1;
    }
    const bool statmc = r.use_statmc != 0;
    INFO("Image: %dx%d", r.image_width, r.image_height);
    INFO("Samples per pixel: %d", r.samples_per_pixel);
    INFO("Camera position: [%g, %g, %g]", st.camera_position[0], st.camera_position[1], st.camera_position[2]);
    INFO("Scene: %s", scene.c_str());
    INFO("Output directory: %s", out_dir.string().c_str());
    INFO("StatMC enabled: %s", statmc ? "true" : "false");
    INFO("Color window radius: %d", r.color_window_radius);
    INFO("Sample-variance window radius: %d", r.var_window_radius);
    INFO("Tonemap: %s", tonemap_name.c_str());

    rpd_context* ctx = nullptr;
    CHECK(rpd_ctx_create(0, &ctx));
    if (statmc && st.debug_statmc_outputs) CHECK(rpd_ctx_set_debug_outputs(ctx, 1));
    rpd_scene* sc = nullptr;
    if (rpd_scene_builtin(ctx, scene.c_str(), &sc) != 0) {
        ERROR("Failed to load scene '%s': %s", scene.c_str(), rpd_last_error());
        rpd_ctx_destroy(ctx);
        return 1;
    }
    int nt = 0, ns = 0, nm = 0, nl = 0;
    rpd_scene_counts(sc, &nt, &ns, &nm, &nl);
    INFO("Loaded scene with %d triangles, %d spheres, %d materials, %d lights", nt, ns, nm, nl);

    uint32_t seed = cli.seed ? *cli.seed : std::random_device{}();
    if (cli.seed) INFO("Using user-provided sampler seed %u", seed);
    else INFO("Using random sampler seed %u", seed);

    // output directory checks and clearing, main.cpp:1877-1911
    out_dir = out_dir.lexically_normal();
    if (out_dir.empty()) { ERROR("Resolved output directory is empty"); return 1; }
    const std::string leaf = out_dir.filename().string();
    if (leaf.empty() || leaf == "." || leaf == "/") {
        ERROR("Output directory '%s' must include a leaf name to derive the PNG filename", out_dir.string().c_str());
        return 1;
    }
    std::error_code ec;
    if (fs::exists(out_dir, ec)) {
        if (!fs::is_directory(out_dir, ec)) { ERROR("Output path '%s' exists but is not a directory", out_dir.string().c_str()); return 1; }
        INFO("Clearing output directory '%s'", out_dir.string().c_str());
        std::error_code cec;
        fs::remove_all(out_dir, cec);
        if (cec) { ERROR("Failed to clear output directory '%s': %s", out_dir.string().c_str(), cec.message().c_str()); return 1; }
    }
    std::error_code mec;
    if (!fs::exists(out_dir) && !fs::create_directories(out_dir, mec) && mec) {
        ERROR("Failed to create output directory '%s': %s", out_dir.string().c_str(), mec.message().c_str());
        return 1;
    }
    const std::string base = (out_dir / leaf).string();

    rpd_camera cam{};
    std::memcpy(cam.position, st.camera_position, 12);
    std::memcpy(cam.direction, st.camera_direction, 12);
    std::memcpy(cam.world_up, st.camera_up, 12);
    cam.fov_degrees = st.fov;
    cam.focus_distance = st.focus_distance;
    cam.aperture = st.aperture;
    INFO(statmc ? "Rendering (StatMC / RPF)..." : "Rendering...");
    rpd_job* job = nullptr;
    CHECK(rpd_job_submit(ctx, sc, &r, &cam, seed, &job));
    CHECK(rpd_job_wait(job));
    double ms[11] = {0};
    rpd_job_stage_ms(job, ms, 11);
    INFO("Rendering completed in %.0f ms", ms[0]);
    if (statmc) {
        if (r.adaptive_passes > 0 && r.adaptive_spp > 0) {
            INFO("Adaptive sample count computation completed in %.0f ms", ms[5]);
            INFO("Adaptive rendering completed in %.0f ms", ms[1]);
        }
        INFO("Sensitivity recompute completed in %.0f ms", ms[2]);
        INFO("StatMC denoise completed in %.0f ms", ms[3]);
        INFO("Sample-variance denoise completed in %.0f ms", ms[4]);
    }

    // output_buffers, core/buffers.h:197-250
    const int W = r.image_width, H = r.image_height;
    const size_t N = size_t(W) * H;
    std::vector<float> rgb(N * 3), sc1(N);
    bool ok = true;
    auto image = [&](int id, const std::string& suffix, bool png) {
        if (rpd_job_read(job, id, rgb.data(), rgb.size() * 4) != 0) { ok = false; return; }
        if (png) ok &= host::write_png_tonemapped(base + suffix + ".png", W, H, rgb.data(), tm);
        ok &= host::write_hdr(base + suffix + ".hdr", W, H, rgb.data());
    };
    auto image_hdr = [&](int id, const std::string& suffix) {
        if (rpd_job_read(job, id, rgb.data(), rgb.size() * 4) != 0) { ok = false; return; }
        ok &= host::write_hdr(base + suffix + ".hdr", W, H, rgb.data());
    };
    auto scalar = [&](int id, const std::string& suffix) {
        if (rpd_job_read(job, id, sc1.data(), sc1.size() * 4) != 0) { ok = false; return; }
        ok &= host::write_hdr_scalar(base + suffix + ".hdr", W, H, sc1.data());
    };
    image(RPD_BUF_RADIANCE, "", true);
    if (statmc) {
        image(RPD_BUF_RAW_RADIANCE, "_raw", true);
        scalar(RPD_BUF_SENS_ALL, "_sensitivity");
        scalar(RPD_BUF_SENS_PIXEL, "_sensitivity_pixel");
        scalar(RPD_BUF_SENS_BRDF, "_sensitivity_brdf");
        scalar(RPD_BUF_SENS_LENS, "_sensitivity_lens");
        scalar(RPD_BUF_SENS_LIGHT, "_sensitivity_light");
        scalar(RPD_BUF_SENS_LIGHT_UV, "_sensitivity_light_uv");
        scalar(RPD_BUF_SENS_LIGHT_SELECT, "_sensitivity_light_select");
        scalar(RPD_BUF_SENS_ENVIRONMENT, "_sensitivity_environment");
        scalar(RPD_BUF_SENS_RR, "_sensitivity_rr");
        scalar(RPD_BUF_SENSITIVITY_CONFIDENCE, "_sensitivity_confidence");
        scalar(RPD_BUF_SENSITIVITY_GRADIENT, "_sensitivity_gradient");
        scalar(RPD_BUF_UNCERTAINTY, "_uncertainty");
        scalar(RPD_BUF_ALPHA, "_alpha");
        scalar(RPD_BUF_VARIANCE_DENOISED, "_vardenoised");
        scalar(RPD_BUF_ADAPTIVE_IMPORTANCE, "_adaptive_importance");
        scalar(RPD_BUF_SAMPLE_FRACTION, "_sample_fraction");
    }
    image_hdr(RPD_BUF_ALBEDO, "_albedo");
    image_hdr(RPD_BUF_NORMAL, "_normal");
    image_hdr(RPD_BUF_WORLD_POS, "_worldpos");
    scalar(RPD_BUF_DEPTH, "_depth");
    if (statmc) {
        scalar(RPD_BUF_VAR_TOTAL, "_var_total");
        scalar(RPD_BUF_VAR_EFF, "_var_eff");
        scalar(RPD_BUF_VAR_RATIO, "_var_ratio");
    }
    // debug_statmc_outputs: written to the working directory, as statmc_denoise
    // (main.cpp:1129-1157) and compute_adaptive_sample_counts (1283-1289,
    // 1319-1321) do on their last call
    if (statmc && st.debug_statmc_outputs) {
        auto debug_scalar = [&](int id, const char* name) {
            if (rpd_job_read(job, id, sc1.data(), sc1.size() * 4) != 0) { ok = false; return; }
            ok &= host::write_hdr_scalar(name, W, H, sc1.data());
        };
        debug_scalar(RPD_BUF_DEBUG_NEIGHBORS, "debug_neighbors.hdr");
        debug_scalar(RPD_BUF_ALPHA, "debug_alpha.hdr");
        debug_scalar(RPD_BUF_VAR_TOTAL, "debug_var_total.hdr");
        debug_scalar(RPD_BUF_VAR_EFF, "debug_var_eff.hdr");
        debug_scalar(RPD_BUF_VAR_RATIO, "debug_var_ratio.hdr");
        debug_scalar(RPD_BUF_DEBUG_WEIGHT_SUM, "debug_weight_sum.hdr");
        debug_scalar(RPD_BUF_DEBUG_NEFF, "debug_neff.hdr");
        debug_scalar(RPD_BUF_SENS_ALL, "debug_sensitivity_all.hdr");
        debug_scalar(RPD_BUF_SENS_PIXEL, "debug_sensitivity_pixel.hdr");
        debug_scalar(RPD_BUF_SENS_BRDF, "debug_sensitivity_brdf.hdr");
        debug_scalar(RPD_BUF_SENS_LENS, "debug_sensitivity_lens.hdr");
        debug_scalar(RPD_BUF_SENS_LIGHT, "debug_sensitivity_light.hdr");
        debug_scalar(RPD_BUF_SENS_LIGHT_UV, "debug_sensitivity_light_uv.hdr");
        debug_scalar(RPD_BUF_SENS_LIGHT_SELECT, "debug_sensitivity_light_select.hdr");
        debug_scalar(RPD_BUF_SENS_ENVIRONMENT, "debug_sensitivity_environment.hdr");
        debug_scalar(RPD_BUF_SENS_RR, "debug_sensitivity_rr.hdr");
        debug_scalar(RPD_BUF_SENSITIVITY_CONFIDENCE, "debug_sensitivity_confidence.hdr");
        debug_scalar(RPD_BUF_SENSITIVITY_GRADIENT, "debug_sensitivity_gradient.hdr");
        // SensitivityState::reliable / reliable_denoise (main.cpp:297-304) of every pixel
        std::vector<float> all(N), grad(N), rel(N), rel_den(N);
        if (rpd_job_read(job, RPD_BUF_SENS_ALL, all.data(), N * 4) != 0 ||
            rpd_job_read(job, RPD_BUF_SENSITIVITY_GRADIENT, grad.data(), N * 4) != 0) {
            ok = false;
        } else {
            for (size_t i = 0; i < N; ++i) {
                const float a = std::min(1.0f, std::max(0.0f, all[i]));
                const float g = std::min(1.0f, std::max(0.0f, grad[i]));
                rel[i] = a;
                rel_den[i] = std::min(1.0f, std::max(0.0f, a * (1.0f - g)));
            }
            ok &= host::write_hdr_scalar("debug_reliability_base.hdr", W, H, rel.data());
            ok &= host::write_hdr_scalar("debug_reliability_denoise.hdr", W, H, rel_den.data());
        }
        if (r.adaptive_passes > 0) {
            // with adaptive_spp 0 the allocation returns its floor of 0 before any pass
            std::vector<int> counts(N, 0);
            if (r.adaptive_spp > 0 && rpd_job_read(job, RPD_BUF_PASS_COUNTS, counts.data(), N * 4) != 0) {
                ok = false;
            } else {
                for (size_t i = 0; i < N; ++i) sc1[i] = float(counts[i]);
                ok &= host::write_hdr_scalar("debug_adaptive_extra_counts.hdr", W, H, sc1.data());
            }
            if (r.adaptive_spp > 0) debug_scalar(RPD_BUF_DEBUG_VARIANCE_SOURCE, "debug_variance_source_adaptive.hdr");
        }
    }
    rpd_job_release(job);
    rpd_scene_destroy(sc);
    rpd_ctx_destroy(ctx);
    if (ok) {
        INFO("Successfully saved output to: %s", out_dir.string().c_str());
        return 0;
    }
    ERROR("Failed to save image");
    return 1;
}
