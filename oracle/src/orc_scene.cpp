// ORACLE — test infrastructure only (see orc_common.h).
#include "orc_scene.h"
#include <algorithm>
#include <numeric>
#include <array>
#include <utility>

namespace orc {

// ---- lights ---------------------------------------------------------------
// core/light.h:22-26
Light make_rect_light(F3 position, F3 width, F3 length, F3 emission) {
    Light l;
    l.kind = LIGHT_RECT;
    l.p0 = position; l.p1 = width; l.p2 = length; l.emission = emission;
    const F3 c = cross(width, length);
    l.area = norm(c);
    l.normal = c / l.area;
    return l;
}
// core/light.h:62-63, core/triangle.h:19-21
Light make_area_light(const Triangle& t) {
    Light l;
    l.kind = LIGHT_AREA;
    l.p0 = t.v0; l.p1 = t.v1; l.p2 = t.v2; l.normal = t.normal; l.emission = t.emission;
    l.area = 0.5f * norm(cross(t.v1 - t.v0, t.v2 - t.v0));
    return l;
}
// core/light.h:47-49, 94-96, 122-124, 141-143
float light_power(const Light& l) {
    if (l.kind == LIGHT_RECT || l.kind == LIGHT_AREA) return l.area * std::max(0.0f, luminance(l.emission));
    return std::max(0.0f, luminance(l.emission));
}

// core/light.h:31-45 (rect), 73-92 (area), 106-120 (point), 137-139 (directional)
LightSample light_sample(const Light& l, F3 p, F2 u) {
    LightSample s;
    s.wi = {0.0f, 0.0f, 0.0f};
    s.dist = 0.0f;
    if (l.kind == LIGHT_DIRECTIONAL) {
        s.wi = l.p0;  // direction stored in p0
        s.dist = std::numeric_limits<float>::infinity();
        return s;
    }
    F3 target;
    if (l.kind == LIGHT_RECT) {
        target = l.p0 + u.x * l.p1 + u.y * l.p2;
    } else if (l.kind == LIGHT_AREA) {
        const float su = std::sqrt(std::max(0.0f, u.x));
        const float b0 = 1.0f - su;
        const float b1 = u.y * su;
        const float b2 = 1.0f - b0 - b1;
        target = b0 * l.p0 + b1 * l.p1 + b2 * l.p2;
    } else {
        target = l.p0;  // point light position
    }
    F3 d = target - p;
    const float sq = sqnorm(d);
    if (sq <= ORC_EPS_SMALL) return s;
    d = normalized(d);
    if (l.kind == LIGHT_RECT || l.kind == LIGHT_AREA) {
        const float cos_light = dot(l.normal, -d);
        if (cos_light <= ORC_EPS_SMALL) return s;
    }
    s.wi = d;
    s.dist = std::sqrt(sq);
    return s;
}

// core/light.h:28-30, 65-67, 126-128, 145-147, 153-163
F3 eval_light(const Light& l, F3 light_dir, float distance) {
    if (l.kind == LIGHT_RECT || l.kind == LIGHT_AREA)
        return (dot(light_dir, l.normal) > 0.0f) ? l.emission : F3{0.0f, 0.0f, 0.0f};
    if (l.kind == LIGHT_POINT) {
        const float falloff = std::max(distance * distance, ORC_EPS_SMALL);
        return l.emission / falloff;
    }
    return l.emission;
}

// core/light.h:51-55, 165-174; core/scene.h:67-69
float light_pdf(const Scene& s, int idx, float select_pdf, F3 /*p*/, F3 dir, float dist) {
    const Light& l = s.lights[idx];
    float pdf = 1.0f;
    if (l.kind == LIGHT_RECT || l.kind == LIGHT_AREA) {
        const float cos_light = dot(l.normal, -dir);
        pdf = (cos_light <= ORC_EPS_SMALL) ? 0.0f : (dist * dist) / (cos_light * l.area);
    }
    return select_pdf * pdf;
}

// core/scene.h:31-65
int Scene::add_light(const Light& l) {
    lights.push_back(l);
    const float p = light_power(l);
    light_powers.push_back(p);
    light_cdf.push_back(light_cdf.empty() ? p : light_cdf.back() + p);
    return int(lights.size()) - 1;
}
float Scene::light_select_pdf(int idx) const {
    const float total = light_cdf.empty() ? 0.0f : light_cdf.back();
    if (total <= ORC_EPS_SMALL) return 0.0f;
    return light_powers[idx] / total;
}
void Scene::sample_light(float u, int& idx, float& pdf) const {
    if (lights.empty()) { idx = -1; pdf = 0.0f; return; }
    const float total = light_cdf.back();
    size_t i = size_t(std::lower_bound(light_cdf.begin(), light_cdf.end(), u * total) - light_cdf.begin());
    if (i >= lights.size()) i = lights.size() - 1;
    idx = int(i);
    pdf = light_select_pdf(idx);
}

// ---- geometry -------------------------------------------------------------
// scenes/scene_utils.h:7-23
Triangle make_triangle(F3 v0, F3 v1, F3 v2, int material_id, F3 emission) {
    Triangle t;
    t.v0 = v0; t.v1 = v1; t.v2 = v2; t.material_id = material_id; t.emission = emission;
    const F3 n = cross(v1 - v0, v2 - v0);
    t.normal = (sqnorm(n) > 0.0f) ? normalized(n) : F3{0.0f, 1.0f, 0.0f};
    return t;
}

// Watertight ray/triangle test, core/triangle.h:35-104 (Woop, Benthin, Wald, JCGT 2013).
bool tri_intersect(const Triangle& tri, const Ray& ray, float t_min, float t_max, float& t_out) {
    const F3 ad{std::fabs(ray.d.x), std::fabs(ray.d.y), std::fabs(ray.d.z)};
    int axis = 0;
    if (ad.y > ad.x && ad.y > ad.z) axis = 1;
    if (ad.z > ad.x && ad.z > ad.y) axis = 2;
    int kz = axis, kx = (kz + 1) % 3, ky = (kx + 1) % 3;
    if (ray.d[kz] < 0.0f) std::swap(kx, ky);
    const float Sx = ray.d[kx] / ray.d[kz];
    const float Sy = ray.d[ky] / ray.d[kz];
    const float Sz = 1.0f / ray.d[kz];
    const F3 A = tri.v0 - ray.o, B = tri.v1 - ray.o, C = tri.v2 - ray.o;
    const float Ax = A[kx] - Sx * A[kz], Ay = A[ky] - Sy * A[kz];
    const float Bx = B[kx] - Sx * B[kz], By = B[ky] - Sy * B[kz];
    const float Cx = C[kx] - Sx * C[kz], Cy = C[ky] - Sy * C[kz];
    float U = Cx * By - Cy * Bx;
    float V = Ax * Cy - Ay * Cx;
    float W = Bx * Ay - By * Ax;
    if (U == 0.0f || V == 0.0f || W == 0.0f) {
        U = float(double(Cx) * double(By) - double(Cy) * double(Bx));
        V = float(double(Ax) * double(Cy) - double(Ay) * double(Cx));
        W = float(double(Bx) * double(Ay) - double(By) * double(Ax));
    }
    if ((U < 0.0f || V < 0.0f || W < 0.0f) && (U > 0.0f || V > 0.0f || W > 0.0f)) return false;
    const float det = U + V + W;
    if (det == 0.0f) return false;
    const float Az = Sz * A[kz], Bz = Sz * B[kz], Cz = Sz * C[kz];
    const float T = U * Az + V * Bz + W * Cz;
    const float rcp = 1.0f / det;
    const float t = T * rcp;
    if (t < t_min || t > t_max) return false;
    t_out = t;
    return true;
}

// core/sphere.h:19-54
bool sphere_intersect(const Sphere& s, const Ray& ray, float t_min, float t_max, float& t_out, F3& n_out) {
    const F3 oc = ray.o - s.center;
    const float a = sqnorm(ray.d);
    const float half_b = dot(oc, ray.d);
    const float c = sqnorm(oc) - s.radius * s.radius;
    const float disc = half_b * half_b - a * c;
    if (disc < 0.0f) return false;
    const float sd = std::sqrt(disc);
    float t = (-half_b - sd) / a;
    if (t < t_min || t > t_max) {
        t = (-half_b + sd) / a;
        if (t < t_min || t > t_max) return false;
    }
    n_out = normalized(ray_at(ray, t) - s.center);
    t_out = t;
    return true;
}

// ---- BVH: core/ray_tracer.h:15-226 -----------------------------------------
static inline float fmin_(float a, float b) { return (b < a) ? b : a; }  // std::min semantics
static inline float fmax_(float a, float b) { return (a < b) ? b : a; }  // std::max semantics

static bool box_intersect(const BvhNode& n, const Ray& r, float& t_near, float& t_far) {
    const F3 inv{1.0f / r.d.x, 1.0f / r.d.y, 1.0f / r.d.z};
    const F3 lo = mul(n.bmin - r.o, inv), hi = mul(n.bmax - r.o, inv);
    const F3 tmn{fmin_(lo.x, hi.x), fmin_(lo.y, hi.y), fmin_(lo.z, hi.z)};
    const F3 tmx{fmax_(lo.x, hi.x), fmax_(lo.y, hi.y), fmax_(lo.z, hi.z)};
    t_near = std::fmax(0.0f, std::fmax(tmn.x, std::fmax(tmn.y, tmn.z)));
    t_far = std::fmin(tmx.x, std::fmin(tmx.y, tmx.z));
    return t_near <= t_far;
}

static int bvh_build_rec(Scene& s, uint32_t first, uint32_t count, int depth) {
    const float inf = std::numeric_limits<float>::infinity();
    F3 bmin{inf, inf, inf}, bmax{-inf, -inf, -inf};
    auto include = [&](F3 p) {
        bmin = {fmin_(bmin.x, p.x), fmin_(bmin.y, p.y), fmin_(bmin.z, p.z)};
        bmax = {fmax_(bmax.x, p.x), fmax_(bmax.y, p.y), fmax_(bmax.z, p.z)};
    };
    for (uint32_t i = 0; i < count; ++i) {
        const Triangle& t = s.triangles[s.tri_index[first + i]];
        include(t.v0); include(t.v1); include(t.v2);
    }
    auto make_leaf = [&]() {
        BvhNode n; n.bmin = bmin; n.bmax = bmax; n.first = first; n.count = count;
        s.nodes.push_back(n);
        return int(s.nodes.size()) - 1;
    };
    if (count <= 4 || depth >= 32) return make_leaf();
    F3 cmin{inf, inf, inf}, cmax{-inf, -inf, -inf};
    for (uint32_t i = 0; i < count; ++i) {
        const F3 c = s.centroids[s.tri_index[first + i]];
        cmin = {fmin_(cmin.x, c.x), fmin_(cmin.y, c.y), fmin_(cmin.z, c.z)};
        cmax = {fmax_(cmax.x, c.x), fmax_(cmax.y, c.y), fmax_(cmax.z, c.z)};
    }
    const F3 ext = cmax - cmin;
    int axis = 0;
    if (ext.y > ext.x) axis = 1;
    if (ext.z > ext[axis]) axis = 2;
    if (ext[axis] <= ORC_EPS) return make_leaf();
    const uint32_t mid = first + count / 2;
    std::nth_element(s.tri_index.begin() + first, s.tri_index.begin() + mid, s.tri_index.begin() + first + count,
                     [&](uint32_t a, uint32_t b) { return s.centroids[a][axis] < s.centroids[b][axis]; });
    if (mid == first || mid == first + count) return make_leaf();
    const int left = bvh_build_rec(s, first, mid - first, depth + 1);
    const int right = bvh_build_rec(s, mid, count - (mid - first), depth + 1);
    BvhNode n; n.bmin = bmin; n.bmax = bmax; n.first = first; n.count = 0; n.left = left; n.right = right;
    s.nodes.push_back(n);
    return int(s.nodes.size()) - 1;
}

void Scene::build_bvh() {
    tri_index.resize(triangles.size());
    centroids.resize(triangles.size());
    std::iota(tri_index.begin(), tri_index.end(), 0u);
    nodes.clear();
    nodes.reserve(triangles.size() * 2);
    for (size_t i = 0; i < triangles.size(); ++i) {
        const Triangle& t = triangles[i];
        centroids[i] = (t.v0 + t.v1 + t.v2) / 3.0f;
    }
    root = triangles.empty() ? -1 : bvh_build_rec(*this, 0, uint32_t(triangles.size()), 0);
}

// core/ray_tracer.h:65-137
static bool bvh_traverse(const Scene& s, const Ray& r, float t_max, bool shadow, float& t_out, uint32_t& tri_out) {
    tri_out = UINT32_MAX;
    t_out = t_max;
    if (s.root < 0 || s.nodes.empty()) return false;
    const float eff = shadow ? t_max - ORC_EPS_ANYHIT : t_max;
    if (eff <= ORC_EPS_SMALL) return false;
    float closest = eff;
    bool any = false;
    std::array<std::pair<int, float>, 34> stack;
    size_t sp = 0;
    float tn, tf;
    if (!box_intersect(s.nodes[s.root], r, tn, tf) || tn > eff || tf < ORC_EPS_SMALL) return false;
    stack[sp++] = {s.root, tn};
    while (sp > 0) {
        const auto [ni, near] = stack[--sp];
        if (near > closest) continue;
        const BvhNode& n = s.nodes[ni];
        if (n.count > 0) {
            for (uint32_t i = 0; i < n.count; ++i) {
                const uint32_t ti = s.tri_index[n.first + i];
                float t;
                if (tri_intersect(s.triangles[ti], r, 0.0f, closest, t) && t < closest) {
                    closest = t; tri_out = ti; any = true;
                    if (shadow) { t_out = closest; return true; }
                }
            }
        } else {
            std::pair<int, float> ent[2];
            int cnt = 0;
            for (int child : {n.left, n.right}) {
                if (child < 0) continue;
                float cn, cf;
                if (!box_intersect(s.nodes[child], r, cn, cf) || cn > closest || cf < ORC_EPS_SMALL) continue;
                ent[cnt++] = {child, cn};
            }
            if (cnt == 1) {
                stack[sp++] = ent[0];
            } else if (cnt == 2) {
                if (ent[0].second < ent[1].second) { stack[sp++] = ent[1]; stack[sp++] = ent[0]; }
                else { stack[sp++] = ent[0]; stack[sp++] = ent[1]; }
            }
        }
    }
    t_out = closest;
    return any;
}

// core/ray_tracer.h:255-344
Hit closest_hit(const Scene& s, const Ray& r) {
    const float inf = std::numeric_limits<float>::infinity();
    float tri_t; uint32_t tri_i;
    const bool tri_hit = bvh_traverse(s, r, inf, false, tri_t, tri_i);
    bool sph_hit = false;
    float sph_t = inf; uint32_t sph_i = UINT32_MAX; F3 sph_n{0.0f, 0.0f, 0.0f};
    for (size_t i = 0; i < s.spheres.size(); ++i) {
        float t; F3 n;
        if (sphere_intersect(s.spheres[i], r, ORC_EPS_SMALL, sph_t, t, n) && t < sph_t) {
            sph_hit = true; sph_t = t; sph_i = uint32_t(i); sph_n = n;
        }
    }
    Hit h;
    if (tri_hit && (!sph_hit || tri_t < sph_t)) {
        h.hit = true; h.t = tri_t; h.is_sphere = 0; h.index = tri_i;
    } else if (sph_hit) {
        h.hit = true; h.t = sph_t; h.is_sphere = 1; h.index = sph_i; h.sphere_normal = sph_n;
    }
    return h;
}

bool any_hit(const Scene& s, const Ray& r, float t_max) {
    float t; uint32_t i;
    if (bvh_traverse(s, r, t_max, true, t, i)) return true;
    const float eff = t_max - ORC_EPS_ANYHIT;
    if (eff <= ORC_EPS_SMALL) return false;
    for (const Sphere& sp : s.spheres) {
        float tt; F3 n;
        if (sphere_intersect(sp, r, ORC_EPS_SMALL, eff, tt, n)) return true;
    }
    return false;
}

Hit closest_hit_brute(const Scene& s, const Ray& r) {
    const float inf = std::numeric_limits<float>::infinity();
    Hit h;
    float best = inf;
    for (size_t i = 0; i < s.triangles.size(); ++i) {
        float t;
        if (tri_intersect(s.triangles[i], r, 0.0f, best, t) && t < best) {
            best = t; h.hit = true; h.t = t; h.is_sphere = 0; h.index = uint32_t(i);
        }
    }
    for (size_t i = 0; i < s.spheres.size(); ++i) {
        float t; F3 n;
        if (sphere_intersect(s.spheres[i], r, ORC_EPS_SMALL, best, t, n) && t < best) {
            best = t; h.hit = true; h.t = t; h.is_sphere = 1; h.index = uint32_t(i); h.sphere_normal = n;
        }
    }
    return h;
}

// ---- materials: core/material.h -------------------------------------------
static const double kPi = 3.14159265358979323846;

// Microfacet pieces, core/material.h:62-117
static F3 mf_fresnel(const Material& m, float c) {
    c = std::clamp(c, 0.0f, 1.0f);
    const F3 dn = m.n1 - m.n2, sn = m.n1 + m.n2;
    const F3 r0 = divv(mul(dn, dn), mul(sn, sn));
    const float f = std::pow(1.0f - c, 5.0f);
    return r0 + (F3{1.0f, 1.0f, 1.0f} - r0) * f;
}
static float mf_D(const Material& m, F3 h) {
    const F3 mm = normalized(h);
    const float c = std::clamp(mm.z, ORC_EPS, 1.0f);
    const float s2 = std::max(0.0f, 1.0f - c * c);
    const float t2 = s2 / (c * c);
    const float a = std::max(m.roughness, ORC_EPS);
    if (m.beckmann) {
        const float rs = a * a;
        return float(double(std::exp(-t2 / rs)) / (kPi * double(rs) * double(std::pow(c, 4.0f))));
    }
    const float den = c * c * (a * a - 1.0f) + 1.0f;
    return float(double(a * a) / (kPi * double(den) * double(den)));
}
static float mf_smith(const Material& m, F3 v) {
    const float c = std::clamp(v.z, ORC_EPS, 1.0f);
    const float s2 = std::max(0.0f, 1.0f - c * c);
    if (s2 <= ORC_EPS_SMALL) return 1.0f;
    const float tn = std::sqrt(s2) / c;
    const float a = std::max(m.roughness, ORC_EPS);
    if (m.beckmann) {
        const float x = 1.0f / (a * tn);
        if (x >= 1.6f) return 1.0f;
        return (3.535f * x + 2.181f * x * x) / (1.0f + 2.276f * x + 2.577f * x * x);
    }
    const float lambda = (-1.0f + std::sqrt(1.0f + a * a * tn * tn)) * 0.5f;
    return 1.0f / (1.0f + lambda);
}

static float mf_pdf(const Material& m, F3 wo_w, F3 wi_w, F3 n) {
    const F3 wo = world_to_local(wo_w, n), wi = world_to_local(wi_w, n);
    if (wo.z <= ORC_EPS_SMALL || wi.z <= ORC_EPS_SMALL) return 0.0f;
    const F3 h = normalized(wo + wi);
    if (!all_finite(h)) return 0.0f;
    const float ndoth = std::max(0.0f, h.z);
    const float wodoth = std::max(0.0f, dot(wo, h));
    if (ndoth <= ORC_EPS_SMALL || wodoth <= ORC_EPS_SMALL) return 0.0f;
    return mf_D(m, h) * ndoth / (4.0f * wodoth);
}

static F3 reflect(F3 d, F3 n) { return d - (2.0f * dot(d, n)) * n; }

F3 brdf_eval(const Material& m, F3 wo_w, F3 wi_w, F3 n) {
    if (m.kind == MAT_LAMBERT) return m.albedo / float(kPi);  // material.h:12-14
    if (m.kind == MAT_MICROFACET) {                             // material.h:119-133
        const F3 wo = normalized(world_to_local(wo_w, n)), wi = normalized(world_to_local(wi_w, n));
        if (wo.z <= ORC_EPS_SMALL || wi.z <= ORC_EPS_SMALL) return {0.0f, 0.0f, 0.0f};
        const F3 h = normalized(wo + wi);
        if (!all_finite(h)) return {0.0f, 0.0f, 0.0f};
        const float co = std::clamp(wo.z, 0.0f, 1.0f), ci = std::clamp(wi.z, 0.0f, 1.0f);
        const float G = mf_smith(m, wo) * mf_smith(m, wi);
        const F3 fr = ((mf_fresnel(m, std::fabs(dot(wi, h))) * G) * mf_D(m, h)) / (4.0f * ci * co);
        return mul(fr, m.albedo);
    }
    // Mirror, material.h:189-199
    const F3 nn = normalized(n);
    if (dot(wo_w, nn) <= ORC_EPS_SMALL || dot(wi_w, nn) <= ORC_EPS_SMALL) return {0.0f, 0.0f, 0.0f};
    const F3 expected = normalized(reflect(-normalized(wo_w), nn));
    const F3 incoming = normalized(wi_w);
    return (sqnorm(expected - incoming) <= ORC_EPS_SMALL) ? m.albedo : F3{0.0f, 0.0f, 0.0f};
}

}  // namespace orc

namespace orc {

// core/material.h:43-46 (Lambertian), 135-150 (Microfacet), 201-208 (Mirror)
float brdf_pdf(const Material& m, F3 wo_w, F3 wi_w, F3 n) {
    if (m.kind == MAT_LAMBERT) {
        const float c = std::max(0.0f, dot(wi_w, n));
        return float(double(c) / kPi);
    }
    if (m.kind == MAT_MICROFACET) return mf_pdf(m, wo_w, wi_w, n);
    const F3 nn = normalized(n);
    if (dot(wo_w, nn) <= ORC_EPS_SMALL || dot(wi_w, nn) <= ORC_EPS_SMALL) return 0.0f;
    const F3 expected = normalized(reflect(-normalized(wo_w), nn));
    const F3 incoming = normalized(wi_w);
    return (sqnorm(expected - incoming) <= ORC_EPS_SMALL) ? 1.0f : 0.0f;
}

// core/material.h:25-37 (Lambertian), 152-182 (Microfacet), 210-220 (Mirror)
void brdf_sample(const Material& m, F3 wo_w, F3 n, F2 u, F3& wi, float& pdf) {
    if (m.kind == MAT_LAMBERT) {
        const float r = std::sqrt(u.x);
        const float theta = float(2.0 * kPi * double(u.y));
        const F3 local{r * std::cos(theta), r * std::sin(theta), std::sqrt(std::max(0.0f, 1.0f - u.x))};
        wi = local_to_world(local, n);
        pdf = float(double(local.z) / kPi);
        return;
    }
    if (m.kind == MAT_MICROFACET) {
        wi = {0.0f, 0.0f, 0.0f}; pdf = 0.0f;
        const F3 wo = normalized(world_to_local(wo_w, n));
        if (wo.z <= ORC_EPS_SMALL) return;
        const float u1 = std::clamp(u.x, ORC_EPS_SMALL, 1.0f - ORC_EPS_SMALL);
        const float u2 = std::clamp(u.y, ORC_EPS_SMALL, 1.0f - ORC_EPS_SMALL);
        const float a = std::max(m.roughness, ORC_EPS);
        F3 wh;
        if (m.beckmann) {
            const float r = a * std::sqrt(-std::log(1.0f - u1));
            const float phi = float(2.0 * kPi * double(u2));
            wh = normalized(F3{-(r * std::cos(phi)), -(r * std::sin(phi)), 1.0f});
        } else {
            const float phi = float(2.0 * kPi * double(u1));
            const float cp = std::cos(phi), sp = std::sin(phi);
            const float ct = std::sqrt((1.0f - u2) / (1.0f + (a * a - 1.0f) * u2));
            const float st = std::sqrt(std::max(0.0f, 1.0f - ct * ct));
            wh = normalized(F3{st * cp, st * sp, ct});
        }
        if (wh.z <= ORC_EPS_SMALL) wh.z = ORC_EPS;
        const F3 wl = normalized(-wo + (2.0f * dot(wo, wh)) * wh);
        if (wl.z <= ORC_EPS_SMALL) return;
        wi = local_to_world(wl, n);
        pdf = mf_pdf(m, wo_w, wi, n);
        return;
    }
    wi = {0.0f, 0.0f, 0.0f}; pdf = 0.0f;
    const F3 nn = normalized(n), wo = normalized(wo_w);
    if (dot(wo, nn) <= ORC_EPS_SMALL) return;
    const F3 r = normalized(reflect(-wo, nn));
    if (!all_finite(r)) return;
    wi = r; pdf = 1.0f;
}

// core/material.h:266-276
F3 brdf_sample_weight(const Material& m, F3 wo, F3 wi, F3 n, float pdf) {
    if (pdf <= ORC_EPS_SMALL) return {0.0f, 0.0f, 0.0f};
    if (brdf_is_delta(m)) return m.albedo;
    return brdf_eval(m, wo, wi, n) * (std::max(0.0f, dot(n, wi)) / pdf);
}

// ---- built-in scene: scenes/cornell_box.h:9-109 ------------------------------
Scene make_cornell_box() {
    Scene s;
    s.environment = {0.0f, 0.0f, 0.0f};
    auto lam = [](float r, float g, float b) { Material m; m.kind = MAT_LAMBERT; m.albedo = {r, g, b}; return m; };
    const int white = s.add_material(lam(0.78f, 0.78f, 0.78f));
    const int red = s.add_material(lam(0.63f, 0.065f, 0.05f));
    const int green = s.add_material(lam(0.14f, 0.45f, 0.091f));
    const int light_mat = s.add_material(lam(1.0f, 1.0f, 1.0f));
    const int box_white = s.add_material(lam(0.73f, 0.73f, 0.73f));
    const float sc = 0.001f;
    auto P = [&](float x, float y, float z) { return F3{x * sc, y * sc, z * sc}; };
    const F3 room_center = P(278.0f, 274.4f, 279.6f);
    const F3 zero{0.0f, 0.0f, 0.0f};
    auto add_oriented = [&](F3 a, F3 b, F3 c, int mat, F3 em, F3 ref, bool toward) {
        Triangle t = make_triangle(a, b, c, mat, em);
        const F3 center = (a + b + c) / 3.0f;
        const float orient = dot(t.normal, ref - center);
        const bool flip = toward ? (orient < 0.0f) : (orient > 0.0f);
        if (flip) t.normal = -t.normal;
        return s.add_triangle(t);
    };
    auto add_quad = [&](F3 a, F3 b, F3 c, F3 d, int mat, F3 ref, bool toward) {
        add_oriented(a, b, c, mat, zero, ref, toward);
        add_oriented(a, c, d, mat, zero, ref, toward);
    };
    const F3 light_color{70.0f, 70.0f, 70.0f};
    const F3 lo = P(213.0f, 548.0f, 227.0f);
    const int light = s.add_light(make_rect_light(lo, P(343.0f, 548.0f, 227.0f) - lo, P(213.0f, 548.0f, 332.0f) - lo, light_color));
    auto add_light_tri = [&](F3 a, F3 b, F3 c) {
        const int t = add_oriented(a, b, c, light_mat, light_color, room_center, true);
        s.triangle_to_light[t] = light;
    };
    const F3 l0 = P(213.0f, 548.0f, 227.0f), l1 = P(343.0f, 548.0f, 227.0f);
    const F3 l2 = P(343.0f, 548.0f, 332.0f), l3 = P(213.0f, 548.0f, 332.0f);
    add_light_tri(l0, l2, l1);
    add_light_tri(l0, l3, l2);
    add_quad(P(552.8f, 0.0f, 0.0f), P(0.0f, 0.0f, 0.0f), P(0.0f, 0.0f, 559.2f), P(549.6f, 0.0f, 559.2f), white, room_center, true);
    add_quad(P(556.0f, 548.8f, 0.0f), P(556.0f, 548.8f, 559.2f), P(0.0f, 548.8f, 559.2f), P(0.0f, 548.8f, 0.0f), white, room_center, true);
    add_quad(P(549.6f, 0.0f, 559.2f), P(0.0f, 0.0f, 559.2f), P(0.0f, 548.8f, 559.2f), P(556.0f, 548.8f, 559.2f), white, room_center, true);
    add_quad(P(0.0f, 0.0f, 559.2f), P(0.0f, 0.0f, 0.0f), P(0.0f, 548.8f, 0.0f), P(0.0f, 548.8f, 559.2f), green, room_center, true);
    add_quad(P(552.8f, 0.0f, 0.0f), P(549.6f, 0.0f, 559.2f), P(556.0f, 548.8f, 559.2f), P(556.0f, 548.8f, 0.0f), red, room_center, true);
    const F3 sb[8] = {P(130.0f, 0.0f, 65.0f), P(82.0f, 0.0f, 225.0f), P(240.0f, 0.0f, 272.0f), P(290.0f, 0.0f, 114.0f),
                      P(130.0f, 165.0f, 65.0f), P(82.0f, 165.0f, 225.0f), P(240.0f, 165.0f, 272.0f), P(290.0f, 165.0f, 114.0f)};
    const F3 sbc = (sb[0] + sb[2] + sb[5] + sb[7]) * 0.25f;
    add_quad(sb[4], sb[5], sb[6], sb[7], box_white, sbc, false);
    add_quad(sb[0], sb[4], sb[7], sb[3], box_white, sbc, false);
    add_quad(sb[3], sb[7], sb[6], sb[2], box_white, sbc, false);
    add_quad(sb[2], sb[6], sb[5], sb[1], box_white, sbc, false);
    add_quad(sb[1], sb[5], sb[4], sb[0], box_white, sbc, false);
    const F3 tb[8] = {P(423.0f, 0.0f, 247.0f), P(265.0f, 0.0f, 296.0f), P(314.0f, 0.0f, 456.0f), P(472.0f, 0.0f, 406.0f),
                      P(423.0f, 330.0f, 247.0f), P(265.0f, 330.0f, 296.0f), P(314.0f, 330.0f, 456.0f), P(472.0f, 330.0f, 406.0f)};
    const F3 tbc = (tb[0] + tb[2] + tb[5] + tb[7]) * 0.25f;
    add_quad(tb[4], tb[5], tb[6], tb[7], box_white, tbc, false);
    add_quad(tb[0], tb[4], tb[7], tb[3], box_white, tbc, false);
    add_quad(tb[3], tb[7], tb[6], tb[2], box_white, tbc, false);
    add_quad(tb[2], tb[6], tb[5], tb[1], box_white, tbc, false);
    add_quad(tb[1], tb[5], tb[4], tb[0], box_white, tbc, false);
    s.build_bvh();
    return s;
}

}  // namespace orc
