// ORACLE — test infrastructure only (see orc_common.h).
// Scene description, materials, lights, intersection and the binary BVH of
// the reference renderer, restated for the CPU oracle.
#pragma once
#include <vector>
#include <cstdint>
#include <limits>
#include "orc_common.h"

namespace orc {

// ---- materials: path_tracer/core/material.h:9-276 -------------------------
enum MatKind : int { MAT_LAMBERT = 0, MAT_MICROFACET = 1, MAT_MIRROR = 2 };
struct Material {
    int kind = MAT_LAMBERT;
    F3 albedo{0.5f, 0.5f, 0.5f};
    float roughness = 0.5f;
    F3 n1{1.0f, 1.0f, 1.0f}, n2{1.5f, 1.5f, 1.5f};
    int beckmann = 0;  // 0 = GGX, 1 = Beckmann
};

// ---- lights: path_tracer/core/light.h:16-178 ------------------------------
enum LightKind : int { LIGHT_RECT = 0, LIGHT_AREA = 1, LIGHT_POINT = 2, LIGHT_DIRECTIONAL = 3 };
struct Light {
    int kind = LIGHT_RECT;
    // rect: position/width/length/normal/emission/area; area: v0,v1,v2,normal,emission,area
    F3 p0, p1, p2, normal, emission;
    float area = 0.0f;
};

struct Triangle {
    F3 v0, v1, v2, normal, emission;
    int material_id = 0;
};
struct Sphere {
    F3 center;
    float radius = 0.0f;
    F3 emission;
    int material_id = 0;
};

inline bool tri_is_emitter(const Triangle& t) { return !is_zero(t.emission, ORC_EPS_SMALL); }
inline bool sphere_is_emitter(const Sphere& s) { return !is_zero(s.emission, ORC_EPS_SMALL); }
inline bool light_is_delta(const Light& l) { return l.kind == LIGHT_POINT || l.kind == LIGHT_DIRECTIONAL; }

Light make_rect_light(F3 position, F3 width, F3 length, F3 emission);
Light make_area_light(const Triangle& t);
float light_power(const Light& l);

struct BvhNode {
    F3 bmin, bmax;
    uint32_t first = 0, count = 0;
    int left = -1, right = -1;
};

struct Scene {
    std::vector<Triangle> triangles;
    std::vector<Sphere> spheres;
    std::vector<Material> materials;
    std::vector<Light> lights;
    std::vector<int> triangle_to_light, sphere_to_light;
    std::vector<float> light_cdf, light_powers;
    F3 environment{0.0f, 0.0f, 0.0f};
    // binary BVH (core/ray_tracer.h:49-226)
    std::vector<BvhNode> nodes;
    std::vector<uint32_t> tri_index;
    std::vector<F3> centroids;
    int root = -1;

    int add_material(const Material& m) { materials.push_back(m); return int(materials.size()) - 1; }
    int add_triangle(const Triangle& t) { triangles.push_back(t); triangle_to_light.push_back(-1); return int(triangles.size()) - 1; }
    int add_sphere(const Sphere& s) { spheres.push_back(s); sphere_to_light.push_back(-1); return int(spheres.size()) - 1; }
    int add_light(const Light& l);
    float light_select_pdf(int idx) const;
    void sample_light(float u, int& idx, float& pdf) const;
    void build_bvh();
};

Triangle make_triangle(F3 v0, F3 v1, F3 v2, int material_id, F3 emission);
Scene make_cornell_box();

struct Ray { F3 o, d; };
inline F3 ray_at(const Ray& r, float t) { return r.o + t * r.d; }

bool tri_intersect(const Triangle& tri, const Ray& ray, float t_min, float t_max, float& t);
bool sphere_intersect(const Sphere& s, const Ray& ray, float t_min, float t_max, float& t, F3& n);

struct Hit {
    bool hit = false;
    float t = 0.0f;
    int is_sphere = 0;
    uint32_t index = UINT32_MAX;
    F3 sphere_normal;
};
Hit closest_hit(const Scene& s, const Ray& r);
bool any_hit(const Scene& s, const Ray& r, float t_max);
// brute force over all primitives (pins the BVH in tests)
Hit closest_hit_brute(const Scene& s, const Ray& r);

// materials
F3 brdf_eval(const Material& m, F3 wo, F3 wi, F3 n);
float brdf_pdf(const Material& m, F3 wo, F3 wi, F3 n);
void brdf_sample(const Material& m, F3 wo, F3 n, F2 u, F3& wi, float& pdf);
F3 brdf_sample_weight(const Material& m, F3 wo, F3 wi, F3 n, float pdf);
inline bool brdf_is_delta(const Material& m) { return m.kind == MAT_MIRROR; }

// lights
struct LightSample { F3 wi; float dist = 0.0f; };
LightSample light_sample(const Light& l, F3 p, F2 u);
F3 eval_light(const Light& l, F3 light_dir, float distance);
float light_pdf(const Scene& s, int idx, float select_pdf, F3 p, F3 dir, float dist);

}  // namespace orc
