// ORACLE — test infrastructure only (see orc_common.h).
// RPF statistics, sensitivity reconstruction, StatMC Welch filter, variance
// smoothing, adaptive allocation and the render() driver of the reference
// (path_tracer/main.cpp:94-231, 741-1564; core/stats.h; core/buffers.h).
#include <chrono>
#include <cstdio>
#include <numeric>
#include "orc_job.h"

namespace orc {

using Clock = std::chrono::high_resolution_clock;
static double ms_since(Clock::time_point t0) {
    return std::chrono::duration<double, std::milli>(Clock::now() - t0).count();
}

// core/stats.h:115-137
template <int N>
float Binned<N>::sensitivity() const {
    constexpr double kEps = 1e-18;
    const double ne = mass();
    if (ne <= 2.0) return 0.0f;
    const double mean = y_sum / w_sum;
    const double total = y2_sum - w_sum * mean * mean;
    if (total <= kEps) return 0.0f;
    int active = 0;
    double between = 0.0;
    for (int b = 0; b < N; ++b) {
        if (weight[b] <= kEps) continue;
        ++active;
        const double d = sum_y[b] / weight[b] - mean;
        between += weight[b] * d * d;
    }
    if (active < 2 || ne <= double(active)) return 0.0f;
    const double eta2 = std::clamp(between / total, 0.0, 1.0);
    const double adj = std::max(0.0, 1.0 - (1.0 - eta2) * (ne - 1.0) / (ne - active));
    return float(std::sqrt(adj));
}
template struct Binned<16>;
template struct Binned<2>;

// core/stats.h:226-240
void Tile::merge(const Tile& o) {
    pixel.b.merge(o.pixel.b); brdf.b.merge(o.brdf.b); screen_brdf.b.merge(o.screen_brdf.b);
    lens.b.merge(o.lens.b); screen_lens.b.merge(o.screen_lens.b);
    light_uv.b.merge(o.light_uv.b); screen_light_uv.b.merge(o.screen_light_uv.b);
    env.b.merge(o.env.b); screen_env.b.merge(o.screen_env.b);
    light_select.merge(o.light_select); screen_light_select.b.merge(o.screen_light_select.b);
    rr.merge(o.rr); screen_rr.b.merge(o.screen_rr.b);
}
// core/stats.h:264-302
void Tile::split(float n0, float sens[TF_COUNT], float conf[TF_COUNT]) const {
    const double n0d = double(n0);
    auto c = [n0d](double m) { return float(m / (m + n0d)); };
    const Binned<16>* b16[TF_COUNT] = {&pixel.b, &brdf.b, &screen_brdf.b, &lens.b, &screen_lens.b, &light_uv.b,
                                       &screen_light_uv.b, &env.b, &screen_env.b, nullptr,
                                       &screen_light_select.b, nullptr, &screen_rr.b};
    for (int f = 0; f < TF_COUNT; ++f) {
        if (f == TF_LIGHT_SELECT) { sens[f] = light_select.sensitivity(); conf[f] = c(light_select.mass()); }
        else if (f == TF_RR) { sens[f] = rr.sensitivity(); conf[f] = c(rr.mass()); }
        else { sens[f] = b16[f]->sensitivity(); conf[f] = c(b16[f]->mass()); }
    }
}
// core/stats.h:312-318
void Grid::init(int w, int h, int size) {
    support = size;
    step = std::max(1, size / 2);
    nx = std::max(2, (w + step - 1) / step + 1);
    ny = std::max(2, (h + step - 1) / step + 1);
    nodes.assign(size_t(nx) * ny, Tile{});
}

// core/buffers.h:121-140
void job_init_buffers(Job& j) {
    const size_t n = size_t(j.w) * j.h;
    const F3 z{0, 0, 0};
    j.radiance.assign(n, z); j.raw_radiance.assign(n, z); j.albedo.assign(n, z); j.normal.assign(n, z);
    j.world_pos.assign(n, z); j.denoised.assign(n, z);
    for (auto* v : {&j.depth, &j.uncertainty, &j.alpha, &j.variance_denoised, &j.var_total, &j.var_eff, &j.var_ratio,
                    &j.sens_confidence, &j.sens_gradient, &j.adaptive_importance})
        v->assign(n, 0.0f);
    j.hit_count.assign(n, 0); j.sample_count.assign(n, 0);
    j.feat_valid = false; j.sens_valid = false;
}

// main.cpp:94-122
void ensure_features(Job& j) {
    if (j.feat_valid) return;
    const int n = j.w * j.h;
    j.feat_normal.assign(n, F3{}); j.feat_albedo.assign(n, F3{}); j.feat_depth.assign(n, 0.0f);
#pragma omp parallel for schedule(static) num_threads(std::max(1, j.nthreads))
    for (int i = 0; i < n; ++i) {
        if (j.hit_count[i] > 0) {
            const float hc = float(j.hit_count[i]);
            F3 nn = j.normal[i] / hc;
            if (sqnorm(nn) > ORC_EPS_SMALL) nn = normalized(nn);
            j.feat_normal[i] = nn;
            j.feat_depth[i] = j.depth[i] / hc;
            j.feat_albedo[i] = j.albedo[i] / hc;
        } else {
            j.feat_normal[i] = j.normal[i]; j.feat_depth[i] = j.depth[i]; j.feat_albedo[i] = j.albedo[i];
        }
    }
    j.feat_valid = true;
}

// core/buffers.h:68-86
float expand_lookup(const Job& j, const std::vector<float>& t, int x, int y) {
    if (j.tiles_step <= 0 || t.empty() || j.tiles_nx == 0 || j.tiles_ny == 0) return 0.0f;
    const float gx = float(x) / float(j.tiles_step), gy = float(y) / float(j.tiles_step);
    const int x0 = std::clamp(int(std::floor(gx)), 0, j.tiles_nx - 1);
    const int y0 = std::clamp(int(std::floor(gy)), 0, j.tiles_ny - 1);
    const int x1 = std::min(x0 + 1, j.tiles_nx - 1), y1 = std::min(y0 + 1, j.tiles_ny - 1);
    const float tx = std::clamp(gx - float(x0), 0.0f, 1.0f), ty = std::clamp(gy - float(y0), 0.0f, 1.0f);
    const int nx = j.tiles_nx;
    const float top = lerpf(t[y0 * nx + x0], t[y0 * nx + x1], tx);
    const float bot = lerpf(t[y1 * nx + x0], t[y1 * nx + x1], tx);
    return lerpf(top, bot, ty);
}
static std::vector<float> expand(const Job& j, const std::vector<float>& t) {
    std::vector<float> out(size_t(j.w) * j.h, 0.0f);
    if (j.tiles_step <= 0 || j.tiles_nx == 0 || t.empty()) return out;
#pragma omp parallel for schedule(static) num_threads(std::max(1, j.nthreads))
    for (int y = 0; y < j.h; ++y)
        for (int x = 0; x < j.w; ++x) out[size_t(y) * j.w + x] = expand_lookup(j, t, x, y);
    return out;
}

// main.cpp:124-154 (only the fields the filters consume are cached)
void ensure_sensitivity_cache(Job& j) {
    if (j.sens_valid) return;
    const int src[6] = {SF_BRDF, SF_LENS, SF_LIGHT_UV, SF_LIGHT_SELECT, SF_ENV, SF_RR};
    for (int k = 0; k < 6; ++k) j.cache_comp[k] = expand(j, j.tiles[src[k]]);
    j.cache_grad = j.sens_gradient;
    if (j.cache_grad.size() != size_t(j.w) * j.h) j.cache_grad.assign(size_t(j.w) * j.h, 0.0f);
    j.sens_valid = true;
}

// core/buffers.h:142-150
static void init_sensitivity_tiles(Job& j) {
    const int size = j.grid.support;
    if (size <= 0 || j.w <= 0 || j.h <= 0) return;
    const int step = std::max(1, size / 2);
    const int nx = std::max(2, (j.w + step - 1) / step + 1), ny = std::max(2, (j.h + step - 1) / step + 1);
    if (size == j.tiles_support && step == j.tiles_step && size_t(nx) * ny == j.tiles[0].size() && !j.tiles[0].empty()) return;
    j.tiles_support = size; j.tiles_step = step; j.tiles_nx = nx; j.tiles_ny = ny;
    for (auto& t : j.tiles) t.assign(size_t(nx) * ny, 0.0f);
}

// main.cpp:156-185
static float local_avg1(const std::vector<float>& v, int w, int h, int cx, int cy) {
    double acc = 0.0, ws = 0.0;
    for (int dy = -1; dy <= 1; ++dy) {
        const int y = std::clamp(cy + dy, 0, h - 1);
        for (int dx = -1; dx <= 1; ++dx) {
            const int x = std::clamp(cx + dx, 0, w - 1);
            const double wt = 1.0 / (1.0 + double(dx * dx + dy * dy));
            acc += wt * double(v[size_t(y) * w + x]);
            ws += wt;
        }
    }
    return ws > 0.0 ? float(acc / ws) : 0.0f;
}
static F3 local_avg3(const std::vector<F3>& v, int w, int h, int cx, int cy) {
    double a0 = 0, a1 = 0, a2 = 0, ws = 0;
    for (int dy = -1; dy <= 1; ++dy) {
        const int y = std::clamp(cy + dy, 0, h - 1);
        for (int dx = -1; dx <= 1; ++dx) {
            const int x = std::clamp(cx + dx, 0, w - 1);
            const double wt = 1.0 / (1.0 + double(dx * dx + dy * dy));
            const F3 q = v[size_t(y) * w + x];
            a0 += double(q.x) * wt; a1 += double(q.y) * wt; a2 += double(q.z) * wt;
            ws += wt;
        }
    }
    if (ws <= 0.0) return {0, 0, 0};
    return {float(a0 / ws), float(a1 / ws), float(a2 / ws)};
}

// main.cpp:741-953
void recompute_sensitivity(Job& j) {
    const auto t0 = Clock::now();
    const int w = j.w, h = j.h, nx = j.grid.nx, ny = j.grid.ny, step = j.grid.step, nn = nx * ny;
    const int nth = std::max(1, j.nthreads);
    j.sens_valid = false;
    init_sensitivity_tiles(j);
    if (j.cfg.rpf_shrinkage_scale <= 0.0f) {
        j.sens_confidence.assign(size_t(w) * h, 0.0f);
        j.sens_gradient.assign(size_t(w) * h, 0.0f);
        ensure_sensitivity_cache(j);
        return;
    }
    ensure_features(j);
    std::vector<float> raw[TF_COUNT], conf[TF_COUNT];
    for (int f = 0; f < TF_COUNT; ++f) { raw[f].assign(nn, 0.0f); conf[f].assign(nn, 0.0f); }
    std::vector<F3> node_n(nn), node_a(nn);
    std::vector<float> node_d(nn);
#pragma omp parallel for schedule(static) num_threads(nth)
    for (int idx = 0; idx < nn; ++idx) {
        const int gy = idx / nx, gx = idx % nx;
        float s[TF_COUNT], c[TF_COUNT];
        j.grid.nodes[idx].split(j.cfg.rpf_confidence_samples, s, c);
        for (int f = 0; f < TF_COUNT; ++f) { raw[f][idx] = s[f]; conf[f][idx] = c[f]; }
        const int px = std::clamp(gx * step, 0, w - 1), py = std::clamp(gy * step, 0, h - 1);
        F3 nrm = local_avg3(j.feat_normal, w, h, px, py);
        if (sqnorm(nrm) > ORC_EPS_SMALL) nrm = normalized(nrm);
        node_n[idx] = nrm;
        node_d[idx] = local_avg1(j.feat_depth, w, h, px, py);
        node_a[idx] = local_avg3(j.feat_albedo, w, h, px, py);
    }
    auto prior_smooth = [&](const std::vector<float>& r, const std::vector<float>& cf) {
        std::vector<float> prior(nn, 0.0f);
#pragma omp parallel for schedule(static) num_threads(nth)
        for (int idx = 0; idx < nn; ++idx) {
            const int gy = idx / nx, gx = idx % nx;
            double ws = 0.0, acc = 0.0;
            for (int dy = -1; dy <= 1; ++dy) {
                const int sy = gy + dy;
                if (sy < 0 || sy >= ny) continue;
                for (int dx = -1; dx <= 1; ++dx) {
                    const int sx = gx + dx;
                    if (sx < 0 || sx >= nx) continue;
                    const int si = sy * nx + sx;
                    const float sw = 1.0f / (1.0f + float(dx * dx + dy * dy));
                    const float nw = std::clamp((dot(node_n[idx], node_n[si]) + 1.0f) * 0.5f, 0.0f, 1.0f);
                    const float dw = std::exp(-4.0f * relative_depth_delta(node_d[idx], node_d[si]));
                    const float aw = std::exp(-2.5f * norm(node_a[idx] - node_a[si]));
                    const double wn = double(sw) * double(nw) * double(dw) * double(aw) * double(std::max(0.05f, cf[si]));
                    acc += wn * double(r[si]);
                    ws += wn;
                }
            }
            prior[idx] = (ws > double(ORC_EPS_SMALL)) ? float(acc / ws) : r[idx];
        }
        return prior;
    };
    std::vector<float> prior[TF_COUNT];
    // same order as the reference: pixel, brdf, lens, light_uv, light_select, env, rr, then 6 screen fields
    const int order[TF_COUNT] = {TF_PIXEL, TF_BRDF, TF_LENS, TF_LIGHT_UV, TF_LIGHT_SELECT, TF_ENV, TF_RR,
                                 TF_SCREEN_BRDF, TF_SCREEN_LENS, TF_SCREEN_LIGHT_UV, TF_SCREEN_LIGHT_SELECT,
                                 TF_SCREEN_ENV, TF_SCREEN_RR};
    for (int k = 0; k < TF_COUNT; ++k) prior[order[k]] = prior_smooth(raw[order[k]], conf[order[k]]);
    // random component k and its screen comparator, in SensitivityState component order
    const int rnd[6] = {TF_BRDF, TF_LENS, TF_LIGHT_UV, TF_LIGHT_SELECT, TF_ENV, TF_RR};
    const int scr[6] = {TF_SCREEN_BRDF, TF_SCREEN_LENS, TF_SCREEN_LIGHT_UV, TF_SCREEN_LIGHT_SELECT, TF_SCREEN_ENV, TF_SCREEN_RR};
    const int sf_val[6] = {SF_BRDF, SF_LENS, SF_LIGHT_UV, SF_LIGHT_SELECT, SF_ENV, SF_RR};
    const int sf_conf[6] = {SF_BRDF_CONF, SF_LENS_CONF, SF_LIGHT_UV_CONF, SF_LIGHT_SELECT_CONF, SF_ENV_CONF, SF_RR_CONF};
#pragma omp parallel for schedule(static) num_threads(nth)
    for (int idx = 0; idx < nn; ++idx) {
        auto fin = [&](int f) { return lerpf(prior[f][idx], raw[f][idx], conf[f][idx]); };
        j.tiles[SF_PIXEL][idx] = std::clamp(fin(TF_PIXEL), 0.0f, 1.0f);
        j.tiles[SF_PIXEL_CONF][idx] = std::clamp(conf[TF_PIXEL][idx], 0.0f, 1.0f);
        float screen[6], rv[6], rc[6], rel[6];
        for (int k = 0; k < 6; ++k) {
            screen[k] = fin(scr[k]);
            rv[k] = std::clamp(fin(rnd[k]), 0.0f, 1.0f);
            rc[k] = std::clamp(conf[rnd[k]][idx], 0.0f, 1.0f);
            j.tiles[sf_conf[k]][idx] = rc[k];
        }
        for (int k = 0; k < 6; ++k) {
            const float relative = rv[k] / std::max(ORC_EPS_SMALL, rv[k] + screen[k]);
            rel[k] = relative * std::min(rc[k], conf[scr[k]][idx]);
        }
        const int dom = int(std::max_element(rel, rel + 6) - rel);
        for (int k = 0; k < 6; ++k) j.tiles[sf_val[k]][idx] = rel[k];
        j.tiles[SF_LIGHT][idx] = std::max(rel[2], rel[3]);
        j.tiles[SF_LIGHT_CONF][idx] = rel[2] >= rel[3] ? rc[2] : rc[3];
        j.tiles[SF_ALL][idx] = rel[dom];
        j.tiles[SF_ALL_CONF][idx] = std::min(rc[dom], conf[scr[dom]][idx]);
    }
    const std::vector<float> es = expand(j, j.tiles[SF_ALL]);
    j.sens_confidence = expand(j, j.tiles[SF_ALL_CONF]);
    j.sens_gradient.assign(size_t(w) * h, 0.0f);
#pragma omp parallel for schedule(static) num_threads(nth)
    for (int y = 0; y < h; ++y) {
        for (int x = 0; x < w; ++x) {
            const float l = es[size_t(y) * w + std::max(0, x - 1)], r = es[size_t(y) * w + std::min(w - 1, x + 1)];
            const float u = es[size_t(std::max(0, y - 1)) * w + x], d = es[size_t(std::min(h - 1, y + 1)) * w + x];
            const float g = std::sqrt((r - l) * (r - l) + (d - u) * (d - u));
            j.sens_gradient[size_t(y) * w + x] = std::clamp(0.5f * g, 0.0f, 1.0f);
        }
    }
    ensure_sensitivity_cache(j);
    j.stage_ms[2] += ms_since(t0);
    if (j.verbose) std::fprintf(stderr, "Sensitivity recompute completed in %.0f ms\n", ms_since(t0));
}

// main.cpp:188-231
float lookup_t_critical(float df, float alpha) {
    df = std::max(df, 1.0f);
    alpha = std::clamp(alpha, ORC_EPS_SMALL, 0.999999f);
    static const float t05[30] = {12.706f, 4.303f, 3.182f, 2.776f, 2.571f, 2.447f, 2.365f, 2.306f, 2.262f, 2.228f,
                                  2.201f, 2.179f, 2.160f, 2.145f, 2.131f, 2.120f, 2.110f, 2.101f, 2.093f, 2.086f,
                                  2.080f, 2.074f, 2.069f, 2.064f, 2.060f, 2.056f, 2.052f, 2.048f, 2.045f, 2.042f};
    static const float t01[30] = {63.657f, 9.925f, 5.841f, 4.604f, 4.032f, 3.707f, 3.499f, 3.355f, 3.250f, 3.169f,
                                  3.106f, 3.055f, 3.012f, 2.977f, 2.947f, 2.921f, 2.898f, 2.878f, 2.861f, 2.845f,
                                  2.831f, 2.819f, 2.807f, 2.797f, 2.787f, 2.779f, 2.771f, 2.763f, 2.756f, 2.750f};
    auto interp = [df](const float* tab) {
        const int lo = std::clamp(int(std::floor(df)), 1, 30), hi = std::clamp(int(std::ceil(df)), 1, 30);
        if (lo == hi) return tab[lo - 1];
        return lerpf(tab[lo - 1], tab[hi - 1], (df - float(lo)) / float(hi - lo));
    };
    if (df < 30.5f && std::fabs(alpha - 0.05f) < ORC_EPS) return interp(t05);
    if (df < 30.5f && std::fabs(alpha - 0.01f) < ORC_EPS) return interp(t01);
    float z;
    if (alpha <= 0.001f) z = 3.29f;
    else if (alpha <= 0.005f) z = 2.81f;
    else if (alpha <= 0.01f) z = 2.58f;
    else if (alpha <= 0.02f) z = 2.33f;
    else if (alpha <= 0.05f) z = 1.96f;
    else if (alpha <= 0.10f) z = 1.64f;
    else z = 1.28f;
    if (df < 30.0f) z *= std::sqrt(df / std::max(1.0f, df - 1.0f));
    return z;
}

// main.cpp:957-1162
void statmc_denoise(Job& j) {
    const auto t0 = Clock::now();
    const int w = j.w, h = j.h, R = j.cfg.color_window_radius;
    const float nth_ = j.cfg.color_normal_threshold, dth = j.cfg.color_depth_threshold, scale = j.cfg.rpf_shrinkage_scale;
    ensure_features(j);
    ensure_sensitivity_cache(j);
    const int n = w * h;
    std::vector<F3> svm(n), svs(n);
    for (int i = 0; i < n; ++i) svm[i] = fmax3(j.stats[i].sqrt_color_variance() / float(std::max(1, j.stats[i].n)), ORC_EPS_SMALL);
    for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x) {
            const int i = y * w + x;
            const F3 a = local_avg3(svm, w, h, x, y);
            svs[i] = {std::max(svm[i].x, a.x), std::max(svm[i].y, a.y), std::max(svm[i].z, a.z)};
        }
    // per-pixel reliable components, main.cpp:297-327
    auto comps = [&](int i, float out[6]) {
        const float g = std::clamp(j.cache_grad[i], 0.0f, 1.0f);
        for (int k = 0; k < 6; ++k) out[k] = std::clamp(std::clamp(j.cache_comp[k][i], 0.0f, 1.0f) * (1.0f - g), 0.0f, 1.0f);
    };
    const float var2_max = j.cfg.color_sigma_max * j.cfg.color_sigma_max;
#pragma omp parallel for schedule(static) num_threads(std::max(1, j.nthreads))
    for (int y = 0; y < h; ++y) {
        for (int x = 0; x < w; ++x) {
            const int i = y * w + x;
            const PixelStats& si = j.stats[i];
            if (j.hit_count[i] == 0) {
                j.denoised[i] = si.color_mean;
                j.uncertainty[i] = si.variance() / float(std::max(1, si.n));
                j.alpha[i] = 1.0f;
                continue;
            }
            const F3 ni = j.feat_normal[i];
            const float di = j.feat_depth[i];
            float var_i = si.variance();
            if (!std::isfinite(var_i)) var_i = ORC_EPS_SMALL;
            var_i = std::clamp(var_i, ORC_EPS_SMALL, var2_max);
            const int n_i = std::max(1, si.n);
            float ci[6]; comps(i, ci);
            const float vm_i = std::max(ORC_EPS_SMALL, var_i / float(n_i));
            j.var_total[i] = var_i;
            j.var_eff[i] = vm_i;
            j.var_ratio[i] = (var_i > ORC_EPS_SMALL) ? (vm_i * float(n_i)) / var_i : 0.0f;
            double ws = 0, wss = 0, wvs = 0, c0 = 0, c1 = 0, c2 = 0, accepted = 0;
            for (int dy = -R; dy <= R; ++dy) {
                const int yy = y + dy;
                if (yy < 0 || yy >= h) continue;
                for (int dx = -R; dx <= R; ++dx) {
                    const int xx = x + dx;
                    if (xx < 0 || xx >= w) continue;
                    const int q = yy * w + xx;
                    if (j.hit_count[q] == 0) continue;
                    if (dot(ni, j.feat_normal[q]) < nth_) continue;
                    if (relative_depth_delta(di, j.feat_depth[q]) > dth) continue;
                    if (norm(j.feat_albedo[i] - j.feat_albedo[q]) > 0.12f) continue;
                    const PixelStats& sq = j.stats[q];
                    float var_j = sq.variance();
                    if (!std::isfinite(var_j)) var_j = ORC_EPS_SMALL;
                    var_j = std::clamp(var_j, ORC_EPS_SMALL, var2_max);
                    const int n_j = std::max(1, sq.n);
                    const float vm_j = std::max(ORC_EPS_SMALL, var_j / float(n_j));
                    double mx = 0.0;
                    for (int ch = 0; ch < 3; ++ch) {
                        const double vi = double(svs[i][ch]), vj = double(svs[q][ch]);
                        const double vs = vi + vj;
                        const double md = double(si.sqrt_color_mean[ch] - sq.sqrt_color_mean[ch]);
                        const double t2 = md * md / std::max(double(ORC_EPS_SMALL), vs);
                        const double den = (vi * vi) / std::max(double(n_i - 1), 1.0) + (vj * vj) / std::max(double(n_j - 1), 1.0);
                        const double df = (den > double(ORC_EPS_SMALL)) ? (vs * vs) / den : double(n_i + n_j - 2);
                        const double tc = double(lookup_t_critical(float(df), j.cfg.color_compat_alpha));
                        mx = std::max(mx, t2 / (tc * tc));
                    }
                    float cj[6]; comps(q, cj);
                    float shared = 0.0f;
                    for (int k = 0; k < 6; ++k) shared = std::max(shared, std::min(ci[k], cj[k]));
                    shared = std::clamp(shared, 0.0f, 1.0f);
                    const float adj = std::clamp(scale * shared, 0.0f, 1.0f);
                    const double qv = mx / (1.0 + double(adj));
                    if (qv > 25.0) continue;
                    const double sp = 1.0 / (1.0 + double(dx * dx + dy * dy));
                    const double wj = sp * std::exp(-0.5 * qv);
                    c0 += double(sq.color_mean.x) * wj; c1 += double(sq.color_mean.y) * wj; c2 += double(sq.color_mean.z) * wj;
                    ws += wj; wss += wj * wj; wvs += wj * wj * double(vm_j);
                    if ((dx != 0 || dy != 0) && wj > 0.05 * sp) accepted += 1.0;
                }
            }
            F3 mn = si.color_mean;
            double vn = double(vm_i);
            if (ws > double(ORC_EPS_SMALL)) {
                mn = {float(c0 / ws), float(c1 / ws), float(c2 / ws)};
                vn = wvs / (ws * ws);
            }
            const double neff = ws * ws / std::max(double(ORC_EPS_SMALL), wss);
            const double blend = (accepted >= 2.0 && ws >= 1.1) ? std::clamp(1.0 - 1.0 / std::sqrt(neff), 0.0, 0.8) : 0.0;
            const float a = float(1.0 - blend);
            j.denoised[i] = lerp3(si.color_mean, mn, float(blend));
            j.uncertainty[i] = float(double(a) * double(a) * double(vm_i) + blend * blend * vn +
                                     2.0 * double(a) * blend * (1.0 / std::max(ws, double(ORC_EPS_SMALL))) * double(vm_i));
            j.alpha[i] = a;
        }
    }
    j.stage_ms[3] += ms_since(t0);
    if (j.verbose) std::fprintf(stderr, "StatMC denoise completed in %.0f ms\n", ms_since(t0));
}

// main.cpp:1166-1268
void variance_denoise(Job& j) {
    const auto t0 = Clock::now();
    const int w = j.w, h = j.h, R = j.cfg.var_window_radius, n = w * h;
    const float maxv = j.cfg.adaptive_sigma_max * j.cfg.adaptive_sigma_max;
    auto rawvar = [&](int i) {
        float v = j.stats[i].variance();
        if (!std::isfinite(v)) v = ORC_EPS_SMALL;
        return std::clamp(v, ORC_EPS_SMALL, maxv);
    };
    std::vector<float> cur(n), nxt(n);
    for (int i = 0; i < n; ++i) cur[i] = rawvar(i);
    ensure_features(j);
    for (int it = 0; it < j.cfg.var_iterations; ++it) {
#pragma omp parallel for schedule(static) num_threads(std::max(1, j.nthreads))
        for (int y = 0; y < h; ++y) {
            for (int x = 0; x < w; ++x) {
                const int i = y * w + x;
                if (j.hit_count[i] == 0) { nxt[i] = cur[i]; continue; }
                const F3 ni = j.feat_normal[i];
                const float di = j.feat_depth[i];
                const float vi = std::max(ORC_EPS_SMALL, cur[i]);
                float ws = 0.0f, wsum = 0.0f;
                for (int dy = -R; dy <= R; ++dy) {
                    const int yy = y + dy;
                    if (yy < 0 || yy >= h) continue;
                    for (int dx = -R; dx <= R; ++dx) {
                        const int xx = x + dx;
                        if (xx < 0 || xx >= w || (dx == 0 && dy == 0)) continue;
                        const int q = yy * w + xx;
                        if (j.hit_count[q] == 0) continue;
                        if (dot(ni, j.feat_normal[q]) < j.cfg.var_normal_threshold) continue;
                        if (relative_depth_delta(di, j.feat_depth[q]) > j.cfg.var_depth_threshold) continue;
                        const float vj = std::max(ORC_EPS_SMALL, cur[q]);
                        const float rd = std::fabs(vi - vj) / std::max(ORC_EPS_SMALL, std::sqrt(vi * vj));
                        if (rd > j.cfg.var_compat_sigma) continue;
                        const float wj = 1.0f / (1.0f + float(dx * dx + dy * dy));
                        wsum += wj * vj;
                        ws += wj;
                    }
                }
                float nm = vi;
                if (ws > ORC_EPS_SMALL) nm = wsum / ws;
                const float blend = std::clamp(std::max(0.0f, nm - vi) / std::max(ORC_EPS_SMALL, vi + nm + j.cfg.var_shrinkage_k), 0.0f, 1.0f);
                nxt[i] = lerpf(vi, nm, blend);
            }
        }
        std::swap(cur, nxt);
    }
    for (int i = 0; i < n; ++i) cur[i] = std::max(cur[i], rawvar(i));
    j.variance_denoised = std::move(cur);
    j.stage_ms[4] += ms_since(t0);
    if (j.verbose) std::fprintf(stderr, "Sample-variance denoise completed in %.0f ms\n", ms_since(t0));
}

// main.cpp:1373-1438: clip-and-redistribute, floors, systematic residual draw.
// `counts` arrives holding the per-pixel floor; `sum` is the sequential FP64
// sum of the importance field.
void counts_from_importance(Job& j, const std::vector<float>& imp, double sum, std::vector<int>& counts) {
    const int n = j.w * j.h;
    const int total = j.cfg.adaptive_spp * j.w * j.h;
    const int floor_pp = std::clamp(j.cfg.adaptive_base_samples, 0, j.cfg.adaptive_spp);
    const int reserved = floor_pp * j.w * j.h;
    int remaining = std::max(0, total - reserved);
    if (sum <= 0.0 || remaining <= 0) return;
    std::vector<float> desired(n, 0.0f);
    std::vector<int> active(n);
    std::iota(active.begin(), active.end(), 0);
    double rem_imp = sum;
    const int max_pp = std::max(4, 8 * std::max(1, j.cfg.adaptive_spp));
    while (!active.empty() && remaining > 0 && rem_imp > double(ORC_EPS_SMALL)) {
        bool clipped = false;
        std::vector<int> next;
        next.reserve(active.size());
        for (int i : active) {
            const double share = double(imp[i]) / rem_imp;
            const double want = share * double(remaining);
            if (want > double(max_pp)) {
                desired[i] = float(max_pp);
                remaining -= max_pp;
                rem_imp -= double(imp[i]);
                clipped = true;
            } else {
                next.push_back(i);
            }
        }
        if (!clipped) {
            for (int i : next) desired[i] = float(double(imp[i]) / rem_imp * double(remaining));
            break;
        }
        active.swap(next);
    }
    int assigned = 0;
    std::vector<float> resid(n, 0.0f);
    double rsum = 0.0;
    for (int i = 0; i < n; ++i) {
        const int add = int(std::floor(desired[i]));
        counts[i] += add;
        assigned += add;
        resid[i] = std::max(0.0f, desired[i] - float(add));
        rsum += double(resid[i]);
    }
    const int left = std::max(0, total - reserved - assigned);
    if (left > 0 && rsum > double(ORC_EPS_SMALL)) {
        const double off = double(hash_to_unit_float(j.seed ^ 0x4d595df4u));
        double cum = 0.0;
        int cur = 0;
        for (int k = 0; k < left; ++k) {
            const double target = ((double(k) + off) / double(left)) * rsum;
            while (cur < n && cum + double(resid[cur]) < target) { cum += double(resid[cur]); ++cur; }
            if (cur < n) counts[cur]++;
        }
    }
}

// main.cpp:1273-1446
void compute_adaptive_counts(Job& j, std::vector<int>& counts) {
    const auto t0 = Clock::now();
    const int w = j.w, h = j.h, n = w * h;
    const int total = j.cfg.adaptive_spp * w * h;
    const int floor_pp = std::clamp(j.cfg.adaptive_base_samples, 0, j.cfg.adaptive_spp);
    const int reserved = floor_pp * w * h;
    counts.assign(n, floor_pp);
    if (total <= 0) return;
    const float smax = j.cfg.adaptive_sigma_max;
    ensure_features(j);
    std::vector<float> src = j.variance_denoised;
    if (src.size() != size_t(n)) src.assign(n, 0.0f);
    for (int i = 0; i < n; ++i) {
        float raw = j.stats[i].variance();
        if (!std::isfinite(raw)) raw = ORC_EPS_SMALL;
        raw = std::max(ORC_EPS_SMALL, raw);
        float d = src[i];
        if (!std::isfinite(d) || d <= 0.0f) d = raw;
        src[i] = std::clamp(std::max(raw, d), ORC_EPS_SMALL, smax * smax);
    }
    std::vector<float> imp(n);
    for (int i = 0; i < n; ++i) imp[i] = std::sqrt(src[i]);
    double sum = std::accumulate(imp.begin(), imp.end(), 0.0);
    const int R = j.cfg.adaptive_importance_smoothing_radius;
    if (R > 0) {
        std::vector<float> sm(n);
#pragma omp parallel for schedule(static) num_threads(std::max(1, j.nthreads))
        for (int y = 0; y < h; ++y) {
            for (int x = 0; x < w; ++x) {
                const int i = y * w + x;
                if (j.hit_count[i] == 0) { sm[i] = imp[i]; continue; }
                double acc = 0.0, ws = 0.0;
                for (int dy = -R; dy <= R; ++dy) {
                    const int yy = y + dy;
                    if (yy < 0 || yy >= h) continue;
                    for (int dx = -R; dx <= R; ++dx) {
                        const int xx = x + dx;
                        if (xx < 0 || xx >= w) continue;
                        const int q = yy * w + xx;
                        if (j.hit_count[q] == 0) continue;
                        const float sw = 1.0f / (1.0f + float(dx * dx + dy * dy));
                        const float nw = std::clamp((dot(j.feat_normal[i], j.feat_normal[q]) + 1.0f) * 0.5f, 0.0f, 1.0f);
                        const float dw = std::exp(-4.0f * relative_depth_delta(j.feat_depth[i], j.feat_depth[q]));
                        const double wn = double(sw) * double(nw) * double(dw);
                        acc += wn * double(imp[q]);
                        ws += wn;
                    }
                }
                sm[i] = (ws > double(ORC_EPS_SMALL)) ? float(acc / ws) : imp[i];
            }
        }
        imp.swap(sm);
        sum = std::accumulate(imp.begin(), imp.end(), 0.0);
    }
    j.adaptive_importance = imp;
    counts_from_importance(j, imp, sum, counts);
    j.stage_ms[5] += ms_since(t0);
    if (j.verbose) std::fprintf(stderr, "Adaptive sample count computation completed in %.0f ms\n", ms_since(t0));
}

// core/buffers.h:152-167
void finalize(Job& j) {
    const int n = j.w * j.h;
    for (int i = 0; i < n; ++i) {
        const float s = float(std::max(1, j.sample_count[i])), hc = float(std::max(1, j.hit_count[i]));
        j.radiance[i] = j.radiance[i] / s;
        j.albedo[i] = j.albedo[i] / hc;
        j.normal[i] = j.normal[i] / hc;
        if (sqnorm(j.normal[i]) > ORC_EPS_SMALL) j.normal[i] = normalized(j.normal[i]);
        j.world_pos[i] = j.world_pos[i] / hc;
        j.depth[i] = j.depth[i] / hc;
    }
}

// main.cpp:1543-1564 (with the StatMC base pass 649-737 and adaptive pass 1451-1538)
void render(Job& j) {
    const auto t_total = Clock::now();
    job_init_buffers(j);
    const int n = j.w * j.h;
    if (!j.cfg.use_statmc) {
        std::vector<int> counts(n, j.cfg.spp);
        const auto t0 = Clock::now();
        render_pass(j, counts.data(), false, false);
        j.stage_ms[0] += ms_since(t0);
        finalize(j);
        j.stage_ms[6] = ms_since(t_total);
        return;
    }
    j.stats.assign(n, PixelStats{});
    j.grid.init(j.w, j.h, j.cfg.rpf_tile_size);
    {
        std::vector<int> counts(n, j.cfg.spp);
        const auto t0 = Clock::now();
        render_pass(j, counts.data(), true, true);
        j.stage_ms[0] += ms_since(t0);
        if (j.verbose) std::fprintf(stderr, "Rendering completed in %.0f ms\n", ms_since(t0));
        const bool dead = j.cfg.adaptive_passes > 0 && j.cfg.adaptive_spp > 0;
        if (!(j.skip_dead_work && dead)) recompute_sensitivity(j);
    }
    for (int p = 0; p < j.cfg.adaptive_passes && j.cfg.adaptive_spp > 0; ++p) {
        if (!j.skip_dead_work) statmc_denoise(j);
        variance_denoise(j);
        compute_adaptive_counts(j, j.extra_counts);
        const auto t0 = Clock::now();
        render_pass(j, j.extra_counts.data(), true, true);
        j.stage_ms[1] += ms_since(t0);
        if (j.verbose) std::fprintf(stderr, "Adaptive rendering completed in %.0f ms\n", ms_since(t0));
        if (!j.skip_dead_work) recompute_sensitivity(j);
    }
    recompute_sensitivity(j);
    statmc_denoise(j);
    variance_denoise(j);
    finalize(j);
    j.raw_radiance = j.radiance;
    j.radiance = j.denoised;
    j.stage_ms[6] = ms_since(t_total);
}

}  // namespace orc
