// ORACLE — test infrastructure only (see orc_common.h).
// Sampler mappings, camera, recursive MIS path tracer and the per-pass sample
// loop of the reference (path_tracer/main.cpp:394-737, 1451-1538).
#include <chrono>
#include <cstdio>
#include "orc_job.h"
#ifdef _OPENMP
#include <omp.h>
#endif

namespace orc {

static const double kPi = 3.14159265358979323846;

// core/sampler.h:22-35 (Shirley concentric disk)
F2 sample_disk(F2 u) {
    const F2 o{2.0f * u.x - 1.0f, 2.0f * u.y - 1.0f};
    if (o.x == 0 && o.y == 0) return {0.0f, 0.0f};
    float theta, r;
    if (std::fabs(o.x) > std::fabs(o.y)) {
        r = o.x;
        theta = float((kPi / 4.0) * double(o.y / o.x));
    } else {
        r = o.y;
        theta = float((kPi / 2.0) - (kPi / 4.0) * double(o.x / o.y));
    }
    return {r * std::cos(theta), r * std::sin(theta)};
}
// core/sampler.h:37-43
F3 sample_sphere(F2 u) {
    const float z = 1.0f - 2.0f * u.x;
    const float r = std::sqrt(std::max(0.0f, 1.0f - z * z));
    const float phi = float(2.0 * kPi * double(u.y));
    return {r * std::cos(phi), r * std::sin(phi), z};
}

// core/camera.h:31-88
void Camera::init(F3 eye, F3 dir, float fov_deg, float aspect_ratio, float focus_dist, float aperture, F3 world_up) {
    position = eye; fov = fov_deg; aspect = aspect_ratio; focus = focus_dist; lens_radius = 0.5f * aperture;
    forward = normalized(dir);
    F3 up_hint = normalized(world_up);
    F3 r = cross(forward, up_hint);
    if (sqnorm(r) < ORC_EPS_SMALL) {
        F3 alt = (std::fabs(forward.z) < 0.999f) ? F3{0, 0, 1} : F3{0, 1, 0};
        const F3 d = alt - up_hint;
        if (sqnorm(d) <= ORC_EPS_SMALL * ORC_EPS_SMALL * std::min(sqnorm(alt), sqnorm(up_hint))) alt = {1, 0, 0};
        up_hint = alt;
        r = cross(forward, up_hint);
    }
    right = normalized(r);
    up = normalized(cross(right, forward));
    const float theta = float(double(fov) * kPi / double(180.0f));
    const float hh = std::tan(0.5f * theta);
    const float hw = aspect * hh;
    horizontal = (2.0f * focus * hw) * right;
    vertical = (2.0f * focus * hh) * up;
    llc = position + forward * focus - 0.5f * horizontal - 0.5f * vertical;
}
Ray Camera::generate(float u, float v, F2 lens) const {
    F3 o = position;
    if (lens_radius > 0.0f) {
        const F2 lo{lens_radius * lens.x, lens_radius * lens.y};
        o = o + (lo.x * right + lo.y * up);
    }
    const F3 px = llc + u * horizontal + v * vertical;
    return {o, normalized(px - o)};
}

// core/stats.h:25-41
void accumulate_sample(PixelStats& s, F3 c) {
    const float lum = luminance(c);
    const float delta = lum - s.mean;
    s.mean += delta / float(++s.n);
    const float delta2 = lum - s.mean;
    s.square_diff += delta * delta2;
    const F3 cd = c - s.color_mean;
    s.color_mean += cd / float(s.n);
    const F3 sc{std::sqrt(std::max(0.0f, c.x)), std::sqrt(std::max(0.0f, c.y)), std::sqrt(std::max(0.0f, c.z))};
    const F3 sd = sc - s.sqrt_color_mean;
    s.sqrt_color_mean += sd / float(s.n);
    s.sqrt_color_square_diff += mul(sd, sc - s.sqrt_color_mean);
}

namespace {

struct Tracer {
    const Scene& sc;
    Sampler& smp;
    RpfSample* rpf;
    int max_depth = 0;  // engine extension: shading points per path, 0 = unbounded

    // main.cpp:394-533
    F3 shade(int is_sphere, uint32_t index, F3 p, F3 n, F3 wo, int depth = 1) {
        const int mid = is_sphere ? sc.spheres[index].material_id : sc.triangles[index].material_id;
        const Material& m = sc.materials[mid];
        const bool delta_brdf = brdf_is_delta(m);
        F3 L_dir{0, 0, 0};

        const float sel_u = smp.next1d();
        int li; float sel_pdf;
        sc.sample_light(sel_u, li, sel_pdf);
        if (rpf && !rpf->light_select_valid && li >= 0 && li < 16) { rpf->light_index = li; rpf->light_select_valid = true; }

        if (li >= 0 && (!delta_brdf || light_is_delta(sc.lights[li]))) {
            const Light& L = sc.lights[li];
            const bool is_delta = light_is_delta(L);
            const F2 lu = smp.next2d();
            if (rpf && !rpf->light_valid && !is_delta) { rpf->light_u = lu; rpf->light_valid = true; }
            const LightSample ls = light_sample(L, p, lu);
            if (ls.dist > ORC_EPS_SMALL || is_delta) {
                const float pdf_light = light_pdf(sc, li, sel_pdf, p, ls.wi, ls.dist);
                if (pdf_light > ORC_EPS_SMALL) {
                    const float pw = std::max(ORC_EPS_SMALL, brdf_pdf(m, wo, ls.wi, n));
                    const float w_light = is_delta ? 1.0f : pdf_light / (pdf_light + pw);
                    const Ray sray{offset_origin(p, n), ls.wi};
                    if (!any_hit(sc, sray, ls.dist)) {
                        const F3 f = brdf_eval(m, wo, ls.wi, n);
                        const F3 em = eval_light(L, -ls.wi, ls.dist);
                        if (max_coeff(em) > 0.0f) {
                            const float g = delta_brdf ? 1.0f : std::max(0.0f, dot(n, ls.wi));
                            const F3 contrib = (mul(em, f) * g) / pdf_light;
                            L_dir += w_light * contrib;
                        }
                    }
                }
            }
        }

        const F2 eu = smp.next2d();
        if (max_coeff(sc.environment) > 0.0f && !delta_brdf) {
            const F3 ed = sample_sphere(eu);
            if (rpf && !rpf->env_valid) { rpf->env_u = eu; rpf->env_valid = true; }
            const float pdf_env = float(1.0 / (4.0 * kPi));
            const float bp = brdf_pdf(m, wo, ed, n);
            const float w_env = pdf_env / (pdf_env + bp);
            const Ray er{offset_origin(p, n), ed};
            if (!any_hit(sc, er, std::numeric_limits<float>::infinity())) {
                const F3 f = brdf_eval(m, wo, ed, n);
                const float c = std::max(0.0f, dot(n, ed));
                L_dir += (w_env * mul(sc.environment, f)) * (c / pdf_env);
            }
        }

        const F2 bu = smp.next2d();
        if (rpf && !delta_brdf && !rpf->brdf_valid) { rpf->brdf_u = bu; rpf->brdf_valid = true; }
        F3 dir; float pdf;
        brdf_sample(m, wo, n, bu, dir, pdf);
        F3 L_ind{0, 0, 0};
        if (pdf > ORC_EPS_SMALL && sqnorm(dir) > 0.0f && (max_depth <= 0 || depth < max_depth)) {
            const Ray br{offset_origin(p, n), dir};
            const Hit h = closest_hit(sc, br);
            bool emitter = false; int ei = -1;
            if (h.hit) {
                if (h.is_sphere) { emitter = sphere_is_emitter(sc.spheres[h.index]); if (emitter) ei = sc.sphere_to_light[h.index]; }
                else { emitter = tri_is_emitter(sc.triangles[h.index]); if (emitter) ei = sc.triangle_to_light[h.index]; }
            }
            if (h.hit && emitter) {
                const Light& E = sc.lights[ei];
                float w_brdf = 1.0f;
                if (!delta_brdf && !light_is_delta(E)) {
                    const float plb = light_pdf(sc, ei, sc.light_select_pdf(ei), p, dir, h.t);
                    if (plb > ORC_EPS_SMALL) w_brdf = pdf / (pdf + plb);
                }
                L_dir += w_brdf * mul(eval_light(E, -dir, h.t), brdf_sample_weight(m, wo, dir, n, pdf));
            } else if (!h.hit) {
                const float pdf_env = float(1.0 / (4.0 * kPi));
                const float w_brdf = delta_brdf ? 1.0f : pdf / (pdf + pdf_env);
                L_dir += w_brdf * mul(sc.environment, brdf_sample_weight(m, wo, dir, n, pdf));
            }
            if (h.hit && !emitter) {
                const float ksi = smp.next1d();
                if (rpf && !rpf->rr_valid) { rpf->rr_survived = ksi < 0.8f; rpf->rr_valid = true; }
                if (ksi < 0.8f) {
                    const F3 np = ray_at(br, h.t);
                    const F3 nn = h.is_sphere ? h.sphere_normal : sc.triangles[h.index].normal;
                    const F3 R = shade(h.is_sphere, h.index, np, nn, -dir, depth + 1);
                    L_ind = mul(R, brdf_sample_weight(m, wo, dir, n, pdf)) / 0.8f;
                }
            }
        }
        return L_dir + L_ind;
    }

    // main.cpp:540-585; primary features go to `rec`
    F3 trace(const Ray& ray, SampleRecord& rec) {
        const Hit h = closest_hit(sc, ray);
        if (!h.hit) return sc.environment;
        const F3 hp = ray_at(ray, h.t);
        F3 n; int mid; bool emitter; int li;
        if (h.is_sphere) {
            const Sphere& s = sc.spheres[h.index];
            n = h.sphere_normal; mid = s.material_id; emitter = sphere_is_emitter(s);
            li = emitter ? sc.sphere_to_light[h.index] : -1;
        } else {
            const Triangle& t = sc.triangles[h.index];
            n = t.normal; mid = t.material_id; emitter = tri_is_emitter(t);
            li = emitter ? sc.triangle_to_light[h.index] : -1;
        }
        rec.hit = 1; rec.normal = n; rec.albedo = sc.materials[mid].albedo; rec.position = hp; rec.depth = h.t;
        if (emitter) return eval_light(sc.lights[li], -ray.d, h.t);
        return shade(h.is_sphere, h.index, hp, n, -normalized(ray.d));
    }
};

// main.cpp:329-376
void accumulate_rpf(Grid& g, int x, int y, const RpfSample& s, float lum) {
    if (!s.pixel_valid && !s.lens_valid && !s.any_shading()) return;
    const float gx = float(x) / float(g.step), gy = float(y) / float(g.step);
    const int x0 = std::clamp(int(std::floor(gx)), 0, g.nx - 1);
    const int y0 = std::clamp(int(std::floor(gy)), 0, g.ny - 1);
    const int x1 = std::min(x0 + 1, g.nx - 1), y1 = std::min(y0 + 1, g.ny - 1);
    const float tx = std::clamp(gx - float(x0), 0.0f, 1.0f), ty = std::clamp(gy - float(y0), 0.0f, 1.0f);
    const struct { int nx, ny; float w; } ent[4] = {
        {x0, y0, (1.0f - tx) * (1.0f - ty)}, {x1, y0, tx * (1.0f - ty)},
        {x0, y1, (1.0f - tx) * ty}, {x1, y1, tx * ty}};
    for (const auto& e : ent) {
        if (e.w <= ORC_EPS_SMALL) continue;
        Tile& t = g.nodes[e.ny * g.nx + e.nx];
        if (s.pixel_valid) {
            const float half = 0.5f * float(g.support);
            const float left = float(e.nx * g.step) - half, top = float(e.ny * g.step) - half;
            const float su = std::clamp((float(x) + s.pixel_u.x - left) / float(g.support), 0.0f, 1.0f);
            const float sv = std::clamp((float(y) + s.pixel_u.y - top) / float(g.support), 0.0f, 1.0f);
            t.pixel.add(su, sv, lum, e.w);
            if (s.brdf_valid) t.screen_brdf.add(su, sv, lum, e.w);
            if (s.lens_valid) t.screen_lens.add(su, sv, lum, e.w);
            if (s.light_valid) t.screen_light_uv.add(su, sv, lum, e.w);
            if (s.light_select_valid) t.screen_light_select.add(su, sv, lum, e.w);
            if (s.env_valid) t.screen_env.add(su, sv, lum, e.w);
            if (s.rr_valid) t.screen_rr.add(su, sv, lum, e.w);
        }
        if (s.lens_valid) t.lens.add(s.lens_u.x, s.lens_u.y, lum, e.w);
        if (s.brdf_valid) t.brdf.add(s.brdf_u.x, s.brdf_u.y, lum, e.w);
        if (s.light_valid) t.light_uv.add(s.light_u.x, s.light_u.y, lum, e.w);
        if (s.env_valid) t.env.add(s.env_u.x, s.env_u.y, lum, e.w);
        if (s.light_select_valid) t.light_select.add(s.light_index, lum, e.w);
        if (s.rr_valid) t.rr.add(s.rr_survived ? 1 : 0, lum, e.w);
    }
}

}  // namespace

// One sampling pass: main.cpp:590-642 (uniform), 649-737 (StatMC base) and
// 1451-1538 (adaptive). `counts[idx]` samples for pixel idx; the sampler of a
// pixel is seeded once with pixel_sample_seed(idx, n) where n is the pixel's
// finite-sample count before the pass (0 for the uniform render).
void render_pass(Job& j, const int* counts, bool statmc, bool seed_with_n) {
    const int w = j.w, h = j.h;
    const bool capture = statmc && j.cfg.rpf_shrinkage_scale > 0.0f;
    const int nthreads = std::max(1, j.nthreads);
    std::vector<Grid> local(capture ? nthreads : 0);
    for (auto& g : local) g.init(w, h, j.cfg.rpf_tile_size);
    std::vector<std::vector<SampleRecord>> recs(j.keep_records ? h : 0);
#pragma omp parallel num_threads(nthreads)
    {
        int tid = 0;
#ifdef _OPENMP
        tid = omp_get_thread_num();
#endif
        Sampler smp;
#pragma omp for schedule(dynamic, 1)
        for (int y = 0; y < h; ++y) {
            for (int x = 0; x < w; ++x) {
                const int idx = y * w + x;
                const int n_before = seed_with_n ? j.stats[idx].n : 0;
                smp.seed(pixel_sample_seed(j.seed, idx, n_before));
                for (int s = 0; s < counts[idx]; ++s) {
                    const F2 pu = smp.next2d();
                    const float u = (float(x) + pu.x) / float(w);
                    const float v = (float(y) + pu.y) / float(h);
                    const F2 lu = smp.next2d();
                    const F2 ls = j.cam.lens_radius > 0.0f ? sample_disk(lu) : F2{0.0f, 0.0f};
                    RpfSample rs;
                    if (capture) {
                        rs.pixel_u = pu; rs.pixel_valid = true;
                        if (j.cam.lens_radius > 0.0f) { rs.lens_u = lu; rs.lens_valid = true; }
                    }
                    const Ray ray = j.cam.generate(u, 1.0f - v, ls);
                    SampleRecord rec;
                    rec.pixel = idx;
                    Tracer tr{j.scene, smp, capture ? &rs : nullptr, j.cfg.max_depth};
                    const F3 L = tr.trace(ray, rec);
                    if (rec.hit) {
                        j.hit_count[idx]++;
                        j.normal[idx] += rec.normal;
                        j.albedo[idx] += rec.albedo;
                        j.world_pos[idx] += rec.position;
                        j.depth[idx] += rec.depth;
                    }
                    rec.radiance = L;
                    rec.finite = all_finite(L) ? 1 : 0;
                    rec.rpf = rs;
                    if (j.keep_records) recs[y].push_back(rec);
                    if (!rec.finite) continue;
                    if (capture) accumulate_rpf(local[tid], x, y, rs, luminance(L));
                    if (statmc) accumulate_sample(j.stats[idx], L);
                    j.radiance[idx] += L;
                    j.sample_count[idx]++;
                }
            }
        }
    }
    if (capture) {
        const int nn = j.grid.nx * j.grid.ny;
#pragma omp parallel for schedule(static) num_threads(nthreads)
        for (int i = 0; i < nn; ++i)
            for (const Grid& g : local) j.grid.nodes[i].merge(g.nodes[i]);
    }
    if (j.keep_records) {
        j.records.clear();
        for (auto& r : recs) j.records.insert(j.records.end(), r.begin(), r.end());
    }
    j.feat_valid = false;
    j.sens_valid = false;
}

// main.cpp:649-737 with the traced sample replaced by a given record: hit
// features for every sample, then only finite samples enter the RPF grid, the
// Welford statistics and the radiance sums. One thread, pixel-major order.
void accumulate_records(Job& j, const int* counts, const SampleRecord* recs) {
    const bool capture = j.cfg.rpf_shrinkage_scale > 0.0f;
    size_t k = 0;
    for (int y = 0; y < j.h; ++y) {
        for (int x = 0; x < j.w; ++x) {
            const int idx = y * j.w + x;
            for (int s = 0; s < counts[idx]; ++s, ++k) {
                const SampleRecord& rec = recs[k];
                if (rec.hit) {
                    j.hit_count[idx]++;
                    j.normal[idx] += rec.normal;
                    j.albedo[idx] += rec.albedo;
                    j.world_pos[idx] += rec.position;
                    j.depth[idx] += rec.depth;
                }
                const F3 L = rec.radiance;
                if (!all_finite(L)) continue;
                if (capture) accumulate_rpf(j.grid, x, y, rec.rpf, luminance(L));
                accumulate_sample(j.stats[idx], L);
                j.radiance[idx] += L;
                j.sample_count[idx]++;
            }
        }
    }
    j.feat_valid = false;
    j.sens_valid = false;
}

}  // namespace orc
