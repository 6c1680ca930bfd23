// ORACLE — test infrastructure only (see orc_common.h).
// Flat C interface for ctypes: whole-pipeline runs, single stages, buffer
// access by name and the reference's small helpers (for known-answer tests).
#include <cstring>
#include <string>
#include "orc_job.h"

using namespace orc;

extern "C" {

// Field order of this struct is the ABI used by tests/oracle_bind.py.
struct orc_config {
    int width, height, spp;
    float fov, focus_distance, aperture;
    float cam_pos[3], cam_dir[3], cam_up[3];
    int use_statmc, rpf_tile_size, color_window_radius;
    float color_normal_threshold, color_depth_threshold, color_compat_alpha, color_sigma_max;
    int var_window_radius;
    float var_normal_threshold, var_depth_threshold, var_compat_sigma, var_shrinkage_k;
    int var_iterations, adaptive_base_samples, adaptive_spp;
    float adaptive_sigma_max;
    int adaptive_passes, adaptive_importance_smoothing_radius;
    float rpf_shrinkage_scale, rpf_confidence_samples;
    int max_depth;
};

// Scene arrays in the memory layout of the engine's scene PODs
// (include/rpd_engine.h): triangles, spheres, materials and lights.
struct orc_tri { float v0[3], v1[3], v2[3], normal[3], emission[3]; int material_id, light_id; };
struct orc_sphere { float center[3], radius, emission[3]; int material_id, light_id; };
struct orc_mat { int kind; float albedo[3], roughness, n1[3], n2[3]; int distribution; };
struct orc_light { int kind; float p0[3], p1[3], p2[3], normal[3], emission[3]; };

static F3 f3(const float* p) { return {p[0], p[1], p[2]}; }

// The same Scene the reference builds from its scene files (core/scene.h:31-65):
// primitives, materials, lights with their power CDF, emitter links, BVH.
static Scene scene_from_arrays(const orc_tri* tris, int nt, const orc_sphere* sph, int ns, const orc_mat* mats,
                               int nm, const orc_light* lights, int nl, const float* env) {
    Scene s;
    for (int i = 0; i < nm; ++i) {
        Material m;
        m.kind = mats[i].kind; m.albedo = f3(mats[i].albedo); m.roughness = mats[i].roughness;
        m.n1 = f3(mats[i].n1); m.n2 = f3(mats[i].n2); m.beckmann = mats[i].distribution == 1 ? 1 : 0;
        s.add_material(m);
    }
    for (int i = 0; i < nl; ++i) {
        const orc_light& L = lights[i];
        Light l;
        if (L.kind == LIGHT_RECT) {
            l = make_rect_light(f3(L.p0), f3(L.p1), f3(L.p2), f3(L.emission));
        } else {
            l.kind = L.kind; l.p0 = f3(L.p0); l.p1 = f3(L.p1); l.p2 = f3(L.p2); l.normal = f3(L.normal);
            l.emission = f3(L.emission);
            if (L.kind == LIGHT_AREA) l.area = 0.5f * norm(cross(l.p1 - l.p0, l.p2 - l.p0));
        }
        s.add_light(l);
    }
    for (int i = 0; i < nt; ++i) {
        Triangle t;
        t.v0 = f3(tris[i].v0); t.v1 = f3(tris[i].v1); t.v2 = f3(tris[i].v2);
        t.normal = f3(tris[i].normal); t.emission = f3(tris[i].emission); t.material_id = tris[i].material_id;
        const int k = s.add_triangle(t);
        if (tri_is_emitter(t)) s.triangle_to_light[k] = tris[i].light_id;
    }
    for (int i = 0; i < ns; ++i) {
        Sphere sp;
        sp.center = f3(sph[i].center); sp.radius = sph[i].radius; sp.emission = f3(sph[i].emission);
        sp.material_id = sph[i].material_id;
        const int k = s.add_sphere(sp);
        if (sphere_is_emitter(sp)) s.sphere_to_light[k] = sph[i].light_id;
    }
    s.environment = f3(env);
    s.build_bvh();
    return s;
}

void* orc_job_create(const orc_config* c, uint32_t seed, int nthreads) {
    Job* j = new Job();
    Config& k = j->cfg;
    k.width = c->width; k.height = c->height; k.spp = c->spp;
    k.fov = c->fov; k.focus_distance = c->focus_distance; k.aperture = c->aperture;
    k.cam_pos = {c->cam_pos[0], c->cam_pos[1], c->cam_pos[2]};
    k.cam_dir = {c->cam_dir[0], c->cam_dir[1], c->cam_dir[2]};
    k.cam_up = {c->cam_up[0], c->cam_up[1], c->cam_up[2]};
    k.use_statmc = c->use_statmc; k.rpf_tile_size = c->rpf_tile_size; k.color_window_radius = c->color_window_radius;
    k.color_normal_threshold = c->color_normal_threshold; k.color_depth_threshold = c->color_depth_threshold;
    k.color_compat_alpha = c->color_compat_alpha; k.color_sigma_max = c->color_sigma_max;
    k.var_window_radius = c->var_window_radius; k.var_normal_threshold = c->var_normal_threshold;
    k.var_depth_threshold = c->var_depth_threshold; k.var_compat_sigma = c->var_compat_sigma;
    k.var_shrinkage_k = c->var_shrinkage_k; k.var_iterations = c->var_iterations;
    k.adaptive_base_samples = c->adaptive_base_samples; k.adaptive_spp = c->adaptive_spp;
    k.adaptive_sigma_max = c->adaptive_sigma_max; k.adaptive_passes = c->adaptive_passes;
    k.adaptive_importance_smoothing_radius = c->adaptive_importance_smoothing_radius;
    k.rpf_shrinkage_scale = c->rpf_shrinkage_scale; k.rpf_confidence_samples = c->rpf_confidence_samples;
    k.max_depth = c->max_depth;
    j->seed = seed;
    j->nthreads = nthreads;
    j->w = k.width; j->h = k.height;
    j->scene = make_cornell_box();
    j->cam.init(k.cam_pos, k.cam_dir, k.fov, float(k.width) / float(k.height), k.focus_distance, k.aperture, k.cam_up);
    job_init_buffers(*j);
    j->stats.assign(size_t(j->w) * j->h, PixelStats{});
    j->grid.init(j->w, j->h, k.rpf_tile_size);
    return j;
}
void orc_job_destroy(void* p) { delete static_cast<Job*>(p); }
// A job on a scene given as arrays instead of the built-in cornell_box.
void* orc_job_create_scene(const orc_config* c, uint32_t seed, int nthreads, const orc_tri* tris, int nt,
                           const orc_sphere* sph, int ns, const orc_mat* mats, int nm, const orc_light* lights,
                           int nl, const float* env) {
    Job* j = static_cast<Job*>(orc_job_create(c, seed, nthreads));
    j->scene = scene_from_arrays(tris, nt, sph, ns, mats, nm, lights, nl, env);
    return j;
}
void orc_job_flags(void* p, int skip_dead, int keep_records, int verbose) {
    Job* j = static_cast<Job*>(p);
    j->skip_dead_work = skip_dead != 0; j->keep_records = keep_records != 0; j->verbose = verbose != 0;
}
void orc_render(void* p) { render(*static_cast<Job*>(p)); }
// one sampling pass (counts per pixel); statmc=1 seeds with the running finite count
void orc_pass(void* p, const int* counts, int statmc) { render_pass(*static_cast<Job*>(p), counts, statmc != 0, statmc != 0); }
void orc_recompute_sensitivity(void* p) { recompute_sensitivity(*static_cast<Job*>(p)); }
// Records in the 32-float layout of orc_get("records"); word 28 holds the
// validity bits (1 pixel, 2 lens, 4 light select, 8 brdf, 16 light, 32 env, 64 rr).
void orc_accumulate_records(void* p, const int* counts, const float* recs, long n) {
    Job* j = static_cast<Job*>(p);
    std::vector<SampleRecord> v(static_cast<size_t>(n));
    for (long i = 0; i < n; ++i) {
        const float* o = recs + i * 32;
        SampleRecord& r = v[static_cast<size_t>(i)];
        r.pixel = int(o[0]);
        r.radiance = {o[1], o[2], o[3]};
        r.finite = int(o[4]); r.hit = int(o[5]);
        r.normal = {o[6], o[7], o[8]}; r.albedo = {o[9], o[10], o[11]}; r.position = {o[12], o[13], o[14]};
        r.depth = o[15];
        RpfSample& s = r.rpf;
        s.pixel_u = {o[16], o[17]}; s.lens_u = {o[18], o[19]}; s.brdf_u = {o[20], o[21]};
        s.light_u = {o[22], o[23]}; s.env_u = {o[24], o[25]};
        s.light_index = int(o[26]); s.rr_survived = o[27] != 0.0f;
        const int f = int(o[28]);
        s.pixel_valid = f & 1; s.lens_valid = f & 2; s.light_select_valid = f & 4; s.brdf_valid = f & 8;
        s.light_valid = f & 16; s.env_valid = f & 32; s.rr_valid = f & 64;
    }
    accumulate_records(*j, counts, v.data());
}
void orc_statmc(void* p) { statmc_denoise(*static_cast<Job*>(p)); }
void orc_variance(void* p) { variance_denoise(*static_cast<Job*>(p)); }
void orc_adaptive_counts(void* p, int* out) {
    Job* j = static_cast<Job*>(p);
    compute_adaptive_counts(*j, j->extra_counts);
    std::memcpy(out, j->extra_counts.data(), j->extra_counts.size() * sizeof(int));
}
// counts from a given importance field (its sum formed as std::accumulate does)
void orc_adaptive_counts_from_importance(void* p, const float* imp, int* out) {
    Job* j = static_cast<Job*>(p);
    const int n = j->w * j->h;
    std::vector<float> v(imp, imp + n);
    double sum = 0.0;
    for (float x : v) sum += double(x);
    const int floor_pp = std::clamp(j->cfg.adaptive_base_samples, 0, j->cfg.adaptive_spp);
    std::vector<int> counts(n, floor_pp);
    if (j->cfg.adaptive_spp * n > 0) counts_from_importance(*j, v, sum, counts);
    std::memcpy(out, counts.data(), size_t(n) * sizeof(int));
}
void orc_finalize(void* p) { finalize(*static_cast<Job*>(p)); }
double orc_stage_ms(void* p, int i) { return static_cast<Job*>(p)->stage_ms[i]; }
void orc_dims(void* p, int* out) {
    Job* j = static_cast<Job*>(p);
    out[0] = j->w; out[1] = j->h; out[2] = j->grid.nx; out[3] = j->grid.ny; out[4] = j->grid.step; out[5] = j->grid.support;
}

// ---- buffers by name ---------------------------------------------------------
static int vec3_buf(Job* j, const std::string& n, std::vector<F3>*& out) {
    if (n == "radiance") out = &j->radiance; else if (n == "raw_radiance") out = &j->raw_radiance;
    else if (n == "albedo") out = &j->albedo; else if (n == "normal") out = &j->normal;
    else if (n == "world_pos") out = &j->world_pos; else if (n == "denoised") out = &j->denoised;
    else return 0;
    return 1;
}
static std::vector<float>* f32_buf(Job* j, const std::string& n) {
    if (n == "depth") return &j->depth; if (n == "uncertainty") return &j->uncertainty;
    if (n == "alpha") return &j->alpha; if (n == "variance_denoised") return &j->variance_denoised;
    if (n == "var_total") return &j->var_total; if (n == "var_eff") return &j->var_eff;
    if (n == "var_ratio") return &j->var_ratio; if (n == "sensitivity_confidence") return &j->sens_confidence;
    if (n == "sensitivity_gradient") return &j->sens_gradient; if (n == "adaptive_importance") return &j->adaptive_importance;
    return nullptr;
}
// Per-pixel statistics as 12 floats: mean, M2, n, color_mean[3], sqrt_mean[3], sqrt_M2[3]
static const int kStatFloats = 12;

long orc_get(void* p, const char* name, void* dst) {
    Job* j = static_cast<Job*>(p);
    const std::string n(name);
    const size_t np = size_t(j->w) * j->h;
    std::vector<F3>* v3;
    if (vec3_buf(j, n, v3)) { std::memcpy(dst, v3->data(), v3->size() * sizeof(F3)); return long(v3->size() * 3); }
    if (auto* f = f32_buf(j, n)) { std::memcpy(dst, f->data(), f->size() * sizeof(float)); return long(f->size()); }
    if (n == "hit_count" || n == "sample_count") {
        const auto& v = n == "hit_count" ? j->hit_count : j->sample_count;
        std::memcpy(dst, v.data(), v.size() * sizeof(int)); return long(v.size());
    }
    if (n == "stats") {
        float* d = static_cast<float*>(dst);
        for (size_t i = 0; i < np; ++i) {
            const PixelStats& s = j->stats[i];
            float* o = d + i * kStatFloats;
            o[0] = s.mean; o[1] = s.square_diff; std::memcpy(&o[2], &s.n, 4);
            o[3] = s.color_mean.x; o[4] = s.color_mean.y; o[5] = s.color_mean.z;
            o[6] = s.sqrt_color_mean.x; o[7] = s.sqrt_color_mean.y; o[8] = s.sqrt_color_mean.z;
            o[9] = s.sqrt_color_square_diff.x; o[10] = s.sqrt_color_square_diff.y; o[11] = s.sqrt_color_square_diff.z;
        }
        return long(np * kStatFloats);
    }
    if (n == "grid") {  // 440 doubles per node, accumulators in TileField order
        double* d = static_cast<double*>(dst);
        for (const Tile& t : j->grid.nodes) {
            auto put = [&](const auto& b) {
                for (double x : b.weight) *d++ = x;
                for (double x : b.sum_y) *d++ = x;
                *d++ = b.w_sum; *d++ = b.w2_sum; *d++ = b.y_sum; *d++ = b.y2_sum;
            };
            put(t.pixel.b); put(t.brdf.b); put(t.screen_brdf.b); put(t.lens.b); put(t.screen_lens.b);
            put(t.light_uv.b); put(t.screen_light_uv.b); put(t.env.b); put(t.screen_env.b);
            put(t.light_select); put(t.screen_light_select.b); put(t.rr); put(t.screen_rr.b);
        }
        return long(j->grid.nodes.size() * 440);
    }
    if (n == "tiles") {  // SF_COUNT arrays of nodes_x*nodes_y floats
        float* d = static_cast<float*>(dst);
        long k = 0;
        for (int f = 0; f < SF_COUNT; ++f) for (float x : j->tiles[f]) d[k++] = x;
        return k;
    }
    if (n == "features") {  // normal_unit[3], depth_avg, albedo_avg[3] per pixel
        ensure_features(*j);
        float* d = static_cast<float*>(dst);
        for (size_t i = 0; i < np; ++i) {
            float* o = d + i * 7;
            o[0] = j->feat_normal[i].x; o[1] = j->feat_normal[i].y; o[2] = j->feat_normal[i].z;
            o[3] = j->feat_depth[i];
            o[4] = j->feat_albedo[i].x; o[5] = j->feat_albedo[i].y; o[6] = j->feat_albedo[i].z;
        }
        return long(np * 7);
    }
    if (n == "records") {  // 32 floats per sample
        float* d = static_cast<float*>(dst);
        for (size_t i = 0; i < j->records.size(); ++i) {
            const SampleRecord& r = j->records[i];
            float* o = d + i * 32;
            std::memset(o, 0, 32 * sizeof(float));
            o[0] = float(r.pixel); o[1] = r.radiance.x; o[2] = r.radiance.y; o[3] = r.radiance.z;
            o[4] = float(r.finite); o[5] = float(r.hit);
            o[6] = r.normal.x; o[7] = r.normal.y; o[8] = r.normal.z;
            o[9] = r.albedo.x; o[10] = r.albedo.y; o[11] = r.albedo.z;
            o[12] = r.position.x; o[13] = r.position.y; o[14] = r.position.z; o[15] = r.depth;
            const RpfSample& s = r.rpf;
            o[16] = s.pixel_u.x; o[17] = s.pixel_u.y; o[18] = s.lens_u.x; o[19] = s.lens_u.y;
            o[20] = s.brdf_u.x; o[21] = s.brdf_u.y; o[22] = s.light_u.x; o[23] = s.light_u.y;
            o[24] = s.env_u.x; o[25] = s.env_u.y; o[26] = float(s.light_index); o[27] = s.rr_survived ? 1.0f : 0.0f;
            o[28] = float((s.pixel_valid ? 1 : 0) | (s.lens_valid ? 2 : 0) | (s.light_select_valid ? 4 : 0) |
                          (s.brdf_valid ? 8 : 0) | (s.light_valid ? 16 : 0) | (s.env_valid ? 32 : 0) | (s.rr_valid ? 64 : 0));
        }
        return long(j->records.size() * 32);
    }
    return -1;
}
long orc_num_records(void* p) { return long(static_cast<Job*>(p)->records.size()); }

long orc_set(void* p, const char* name, const void* src) {
    Job* j = static_cast<Job*>(p);
    const std::string n(name);
    const size_t np = size_t(j->w) * j->h;
    std::vector<F3>* v3;
    if (vec3_buf(j, n, v3)) { std::memcpy(v3->data(), src, v3->size() * sizeof(F3)); j->feat_valid = false; return long(np * 3); }
    if (auto* f = f32_buf(j, n)) {
        f->resize(np);
        std::memcpy(f->data(), src, np * sizeof(float));
        j->feat_valid = false; j->sens_valid = false;
        return long(np);
    }
    if (n == "hit_count" || n == "sample_count") {
        auto& v = n == "hit_count" ? j->hit_count : j->sample_count;
        std::memcpy(v.data(), src, np * sizeof(int)); j->feat_valid = false; return long(np);
    }
    if (n == "stats") {
        const float* d = static_cast<const float*>(src);
        for (size_t i = 0; i < np; ++i) {
            PixelStats& s = j->stats[i];
            const float* o = d + i * kStatFloats;
            s.mean = o[0]; s.square_diff = o[1]; std::memcpy(&s.n, &o[2], 4);
            s.color_mean = {o[3], o[4], o[5]}; s.sqrt_color_mean = {o[6], o[7], o[8]};
            s.sqrt_color_square_diff = {o[9], o[10], o[11]};
        }
        return long(np * kStatFloats);
    }
    if (n == "grid") {
        const double* d = static_cast<const double*>(src);
        for (Tile& t : j->grid.nodes) {
            auto get = [&](auto& b) {
                for (double& x : b.weight) x = *d++;
                for (double& x : b.sum_y) x = *d++;
                b.w_sum = *d++; b.w2_sum = *d++; b.y_sum = *d++; b.y2_sum = *d++;
            };
            get(t.pixel.b); get(t.brdf.b); get(t.screen_brdf.b); get(t.lens.b); get(t.screen_lens.b);
            get(t.light_uv.b); get(t.screen_light_uv.b); get(t.env.b); get(t.screen_env.b);
            get(t.light_select); get(t.screen_light_select.b); get(t.rr); get(t.screen_rr.b);
        }
        j->sens_valid = false;
        return long(j->grid.nodes.size() * 440);
    }
    return -1;
}

// ---- helpers restated from the reference (known-answer tests) ---------------
uint32_t orc_hash_u32(uint32_t x) { return hash_u32(x); }
float orc_hash_to_unit_float(uint32_t x) { return hash_to_unit_float(x); }
uint32_t orc_pixel_sample_seed(uint32_t base, int idx, int n) { return pixel_sample_seed(base, idx, n); }
void orc_sampler_draws(uint32_t seed, int n, float* out) {
    Sampler s; s.seed(seed);
    for (int i = 0; i < n; ++i) out[i] = s.next1d();
}
void orc_sampler_pairs(uint32_t seed, int n, float* out) {
    Sampler s; s.seed(seed);
    for (int i = 0; i < n; ++i) { const F2 p = s.next2d(); out[2 * i] = p.x; out[2 * i + 1] = p.y; }
}
uint32_t orc_mt_u32(uint32_t seed, int k) {  // k-th (1-based) raw mt19937 output
    std::mt19937 r(seed); uint32_t v = 0;
    for (int i = 0; i < k; ++i) v = uint32_t(r());
    return v;
}
float orc_t_critical(float df, float alpha) { return lookup_t_critical(df, alpha); }
void orc_sample_disk(float ux, float uy, float* out) { const F2 d = sample_disk({ux, uy}); out[0] = d.x; out[1] = d.y; }
void orc_sample_sphere(float ux, float uy, float* out) { const F3 d = sample_sphere({ux, uy}); out[0] = d.x; out[1] = d.y; out[2] = d.z; }
void orc_accumulate_samples(const float* rgb, int n, float* stat12) {
    PixelStats s;
    for (int i = 0; i < n; ++i) accumulate_sample(s, {rgb[3 * i], rgb[3 * i + 1], rgb[3 * i + 2]});
    stat12[0] = s.mean; stat12[1] = s.square_diff; stat12[2] = float(s.n);
    stat12[3] = s.color_mean.x; stat12[4] = s.color_mean.y; stat12[5] = s.color_mean.z;
    const F3 sv = s.sqrt_color_variance();
    stat12[6] = sv.x; stat12[7] = sv.y; stat12[8] = sv.z;
    stat12[9] = s.sqrt_color_mean.x; stat12[10] = s.sqrt_color_mean.y; stat12[11] = s.sqrt_color_mean.z;
}
// nbins 16 or 2 (BinnedDependency<N>); out = {sensitivity, mass}
void orc_binned(int nbins, const int* bins, const float* y, const float* w, int n, double* out) {
    if (nbins == 2) {
        Binned<2> b; for (int i = 0; i < n; ++i) b.add(bins[i], y[i], w[i]);
        out[0] = b.sensitivity(); out[1] = b.mass();
    } else {
        Binned<16> b; for (int i = 0; i < n; ++i) b.add(bins[i], y[i], w[i]);
        out[0] = b.sensitivity(); out[1] = b.mass();
    }
}
// Dependency2D<4>::add(u, v, y, w=1) over n samples; returns sensitivity
float orc_dep2d(const float* u, const float* v, const float* y, int n) {
    Dep2D d; for (int i = 0; i < n; ++i) d.add(u[i], v[i], y[i], 1.0f);
    return d.b.sensitivity();
}
// confidence m/(m+n0), core/stats.h:282-286
float orc_confidence(double mass, float n0) { return float(mass / (mass + double(n0))); }
// core/stats.h:48-91
float orc_relative_random_sensitivity(float r, float s) { return r / std::max(ORC_EPS_SMALL, r + s); }
float orc_shared_reliability(const float* a, const float* b) {
    float sh = 0.0f; for (int k = 0; k < 6; ++k) sh = std::max(sh, std::min(a[k], b[k]));
    return std::clamp(sh, 0.0f, 1.0f);
}
double orc_relax_compatibility(double score, float rel, float scale) {
    const float adj = std::clamp(scale * rel, 0.0f, 1.0f);
    return score / (1.0 + double(adj));
}
float orc_conservative_variance_update(float v, float nb, float k) {
    const float b = std::clamp(std::max(0.0f, nb - v) / std::max(ORC_EPS_SMALL, v + nb + k), 0.0f, 1.0f);
    return lerpf(v, nb, b);
}
// material sampling / evaluation for closed-form checks; kind 0 lambert, 1 GGX, 2 mirror, 3 beckmann
static Material mk(int kind, const float* albedo, float rough) {
    Material m; m.kind = kind == 3 ? MAT_MICROFACET : kind; m.beckmann = kind == 3;
    m.albedo = {albedo[0], albedo[1], albedo[2]}; m.roughness = rough; return m;
}
void orc_brdf_sample(int kind, const float* albedo, float rough, const float* wo, const float* n, float u0, float u1, float* out) {
    const Material m = mk(kind, albedo, rough);
    F3 wi; float pdf;
    brdf_sample(m, {wo[0], wo[1], wo[2]}, {n[0], n[1], n[2]}, {u0, u1}, wi, pdf);
    const F3 bw = brdf_sample_weight(m, {wo[0], wo[1], wo[2]}, wi, {n[0], n[1], n[2]}, pdf);
    out[0] = wi.x; out[1] = wi.y; out[2] = wi.z; out[3] = pdf; out[4] = bw.x; out[5] = bw.y; out[6] = bw.z;
}
void orc_brdf_eval(int kind, const float* albedo, float rough, const float* wo, const float* wi, const float* n, float* out) {
    const Material m = mk(kind, albedo, rough);
    const F3 f = brdf_eval(m, {wo[0], wo[1], wo[2]}, {wi[0], wi[1], wi[2]}, {n[0], n[1], n[2]});
    out[0] = f.x; out[1] = f.y; out[2] = f.z;
    out[3] = brdf_pdf(m, {wo[0], wo[1], wo[2]}, {wi[0], wi[1], wi[2]}, {n[0], n[1], n[2]});
}
// cornell scene triangles: 3 verts + normal + emission + material per triangle (16 floats)
long orc_cornell_triangles(float* out) {
    const Scene s = make_cornell_box();
    for (size_t i = 0; i < s.triangles.size(); ++i) {
        const Triangle& t = s.triangles[i];
        float* o = out + i * 16;
        const F3 v[5] = {t.v0, t.v1, t.v2, t.normal, t.emission};
        for (int k = 0; k < 5; ++k) { o[3 * k] = v[k].x; o[3 * k + 1] = v[k].y; o[3 * k + 2] = v[k].z; }
        o[15] = float(t.material_id);
    }
    return long(s.triangles.size());
}
// closest hit through the BVH and by brute force; out = {hit, t, index} for each
void orc_cornell_hits(const float* rays, int n, float* out_bvh, float* out_brute) {
    static const Scene s = make_cornell_box();
    for (int i = 0; i < n; ++i) {
        const Ray r{{rays[6 * i], rays[6 * i + 1], rays[6 * i + 2]}, {rays[6 * i + 3], rays[6 * i + 4], rays[6 * i + 5]}};
        const Hit a = closest_hit(s, r), b = closest_hit_brute(s, r);
        out_bvh[3 * i] = a.hit; out_bvh[3 * i + 1] = a.t; out_bvh[3 * i + 2] = a.hit ? float(a.index) : -1.0f;
        out_brute[3 * i] = b.hit; out_brute[3 * i + 1] = b.t; out_brute[3 * i + 2] = b.hit ? float(b.index) : -1.0f;
    }
}

}  // extern "C"
