// Host-side scene construction shared by the engine and the scene library:
// a float3 with the evaluation order of the reference's Eigen expressions
// (dot products as x0 + (x1 + x2), normalisation as division by the square
// root), and the arrays of a scene in the engine's ABI form.
#pragma once
#include <cmath>
#include <string>
#include <vector>

#include "../../include/rpd_engine.h"

namespace rpd_host {

struct h3 { float x, y, z; };
inline h3 H3(float a, float b, float c) { return {a, b, c}; }
inline h3 operator+(h3 a, h3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline h3 operator-(h3 a, h3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline h3 operator-(h3 a) { return {-a.x, -a.y, -a.z}; }
inline h3 operator*(float s, h3 a) { return {s * a.x, s * a.y, s * a.z}; }
inline h3 operator*(h3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
inline h3 operator/(h3 a, float s) { return {a.x / s, a.y / s, a.z / s}; }
inline float hdot(h3 a, h3 b) { return a.x * b.x + (a.y * b.y + a.z * b.z); }
inline float hsq(h3 a) { return hdot(a, a); }
inline h3 hnorm(h3 a) {
    const float z = hsq(a);
    if (z > 0.0f) { const float s = std::sqrt(z); return {a.x / s, a.y / s, a.z / s}; }
    return a;
}
inline h3 hcross(h3 a, h3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
inline void put3(float* d, h3 v) { d[0] = v.x; d[1] = v.y; d[2] = v.z; }
inline h3 get3(const float* s) { return {s[0], s[1], s[2]}; }

struct HostScene {
    std::vector<rpd_triangle> tris;
    std::vector<rpd_sphere> spheres;
    std::vector<rpd_material> mats;
    std::vector<rpd_light> lights;
    float env[3] = {0, 0, 0};
};

// The reference's scene registry (scenes/scene_registry.h:17-53): fills `s`
// with the named scene; mesh scenes read resources/models/<mesh>.obj relative
// to the working directory, as the reference does. Returns RPD_OK or an
// RPD_ERR_* code with the message in `err`.
int build_builtin_scene(const std::string& name, HostScene& s, std::string& err);

// Triangles of an OBJ file (core/obj_loader.h:19-65): vertex positions of
// every face, polygons split into a fan from their first vertex, normal from
// the winding, no emission, the given material.
int load_obj_triangles(const std::string& path, int material_id, std::vector<rpd_triangle>& out, std::string& err);

}  // namespace rpd_host
