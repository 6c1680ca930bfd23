// The reference's built-in scenes (path_tracer/scenes/*.h, registry in
// scenes/scene_registry.h:17-53) as arrays of the engine's scene PODs, and a
// reader for the OBJ meshes two of them load (core/obj_loader.h:19-65).
// Coordinates, materials and lights are the reference's scene data; vertex
// arithmetic follows its Eigen expressions in float (sums left to right, dot
// products x0 + (x1 + x2), normalisation as division by the square root).
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>

#include "host_scene.h"

namespace rpd_host {

namespace {

constexpr float kEpsSmall = 1e-8f;  // shared/global.h:4

// ---- scene assembly (core/scene.h:31-65, scenes/scene_utils.h:7-44) --------------
int add_material(HostScene& s, const rpd_material& m) { s.mats.push_back(m); return int(s.mats.size()) - 1; }
int add_light(HostScene& s, const rpd_light& l) { s.lights.push_back(l); return int(s.lights.size()) - 1; }
int add_triangle(HostScene& s, const rpd_triangle& t) { s.tris.push_back(t); return int(s.tris.size()) - 1; }
void link_triangle_to_light(HostScene& s, int tri, int light) { s.tris[size_t(tri)].light_id = light; }

rpd_material lambertian(float r, float g, float b) {
    rpd_material m{};
    m.kind = RPD_MAT_LAMBERTIAN; m.albedo[0] = r; m.albedo[1] = g; m.albedo[2] = b;
    m.roughness = 0.5f;
    return m;
}
rpd_material microfacet(h3 albedo, float roughness, h3 n2) {
    rpd_material m{};
    m.kind = RPD_MAT_MICROFACET;
    put3(m.albedo, albedo);
    m.roughness = roughness;
    put3(m.n1, H3(1.0f, 1.0f, 1.0f));
    put3(m.n2, n2);
    m.distribution = RPD_MICROFACET_GGX;
    return m;
}
rpd_material mirror(h3 albedo) {
    rpd_material m{};
    m.kind = RPD_MAT_MIRROR;
    put3(m.albedo, albedo);
    m.roughness = 0.5f;
    return m;
}

// scenes/material_library.h:7-57
rpd_material lib_gold() { return microfacet(H3(1.0f, 0.71f, 0.29f), 0.0005f, H3(0.17f, 0.35f, 1.5f)); }
rpd_material lib_jade() { return microfacet(H3(0.35f, 0.8f, 0.55f), 0.05f, H3(1.55f, 1.55f, 1.55f)); }
rpd_material lib_mirror_white() { return mirror(H3(0.95f, 0.95f, 0.95f)); }
rpd_material lib_copper() { return microfacet(H3(0.95f, 0.64f, 0.54f), 0.1f, H3(0.27f, 0.67f, 1.38f)); }
rpd_material lib_aluminum() { return microfacet(H3(0.91f, 0.92f, 0.92f), 0.05f, H3(1.44f, 0.96f, 0.62f)); }
rpd_material lib_silver() { return microfacet(H3(0.97f, 0.96f, 0.91f), 0.02f, H3(0.16f, 0.14f, 0.13f)); }

// scenes/scene_utils.h:7-23: normal from the winding, +y when degenerate
rpd_triangle make_triangle(h3 v0, h3 v1, h3 v2, int material, h3 emission = H3(0, 0, 0)) {
    rpd_triangle t{};
    put3(t.v0, v0); put3(t.v1, v1); put3(t.v2, v2);
    put3(t.emission, emission);
    const h3 n = hcross(v1 - v0, v2 - v0);
    put3(t.normal, hsq(n) > 0.0f ? hnorm(n) : H3(0.0f, 1.0f, 0.0f));
    t.material_id = material;
    t.light_id = -1;
    return t;
}
rpd_light rect_light(h3 pos, h3 width, h3 height, h3 emission) {
    rpd_light l{};
    l.kind = RPD_LIGHT_RECT;
    put3(l.p0, pos); put3(l.p1, width); put3(l.p2, height); put3(l.emission, emission);
    return l;
}
rpd_light point_light(h3 pos, h3 intensity) {
    rpd_light l{};
    l.kind = RPD_LIGHT_POINT;
    put3(l.p0, pos); put3(l.emission, intensity);
    return l;
}
rpd_sphere sphere(h3 c, float r, int material) {
    rpd_sphere sp{};
    put3(sp.center, c);
    sp.radius = r;
    sp.material_id = material;
    sp.light_id = -1;
    return sp;
}
// scenes/scene_utils.h:31-44
void add_ground_plane(HostScene& s, float size, float y, int material) {
    const float half = size * 0.5f;
    const h3 v0 = H3(-half, y, -half), v1 = H3(half, y, -half), v2 = H3(half, y, half), v3 = H3(-half, y, half);
    add_triangle(s, make_triangle(v0, v2, v1, material));
    add_triangle(s, make_triangle(v0, v3, v2, material));
}
// a rect light and its two emitting triangles (o, o+u, o+u+v) and (o, o+u+v, o+v)
void add_rect_light_quad(HostScene& s, h3 o, h3 u, h3 v, h3 color, int light_mat) {
    const int li = add_light(s, rect_light(o, u, v, color));
    const int t1 = add_triangle(s, make_triangle(o, o + u, o + u + v, light_mat, color));
    const int t2 = add_triangle(s, make_triangle(o, o + u + v, o + v, light_mat, color));
    link_triangle_to_light(s, t1, li);
    link_triangle_to_light(s, t2, li);
}

// ---- scenes/cornell_box.h:9-109 -----------------------------------------------------
void build_cornell(HostScene& s) {
    auto lam = [&](float r, float g, float b) { return add_material(s, lambertian(r, g, b)); };
    const int white = lam(0.78f, 0.78f, 0.78f), red = lam(0.63f, 0.065f, 0.05f), green = lam(0.14f, 0.45f, 0.091f);
    const int light_mat = lam(1.0f, 1.0f, 1.0f), box_white = lam(0.73f, 0.73f, 0.73f);
    const float sc = 0.001f;
    auto P = [&](float x, float y, float z) { return H3(x * sc, y * sc, z * sc); };
    const h3 center = P(278.0f, 274.4f, 279.6f);
    const h3 zero = H3(0, 0, 0);
    auto add_tri = [&](h3 a, h3 b, h3 c, int mat, h3 em, h3 ref, bool toward, int light) {
        rpd_triangle t{};
        put3(t.v0, a); put3(t.v1, b); put3(t.v2, c); put3(t.emission, em);
        const h3 n = hcross(b - a, c - a);
        h3 nn = hsq(n) > 0.0f ? hnorm(n) : H3(0, 1, 0);
        const h3 mid = (a + b + c) / 3.0f;
        const float orient = hdot(nn, ref - mid);
        if (toward ? (orient < 0.0f) : (orient > 0.0f)) nn = -nn;
        put3(t.normal, nn);
        t.material_id = mat;
        t.light_id = light;
        s.tris.push_back(t);
    };
    auto quad = [&](h3 a, h3 b, h3 c, h3 d, int mat, h3 ref, bool toward) {
        add_tri(a, b, c, mat, zero, ref, toward, -1);
        add_tri(a, c, d, mat, zero, ref, toward, -1);
    };
    const h3 em = H3(70.0f, 70.0f, 70.0f);
    const h3 lo = P(213.0f, 548.0f, 227.0f);
    rpd_light L{};
    L.kind = RPD_LIGHT_RECT;
    put3(L.p0, lo); put3(L.p1, P(343.0f, 548.0f, 227.0f) - lo); put3(L.p2, P(213.0f, 548.0f, 332.0f) - lo);
    put3(L.emission, em);
    s.lights.push_back(L);
    const h3 l0 = P(213.0f, 548.0f, 227.0f), l1 = P(343.0f, 548.0f, 227.0f), l2 = P(343.0f, 548.0f, 332.0f),
             l3 = P(213.0f, 548.0f, 332.0f);
    add_tri(l0, l2, l1, light_mat, em, center, true, 0);
    add_tri(l0, l3, l2, light_mat, em, center, true, 0);
    quad(P(552.8f, 0, 0), P(0, 0, 0), P(0, 0, 559.2f), P(549.6f, 0, 559.2f), white, center, true);
    quad(P(556.0f, 548.8f, 0), P(556.0f, 548.8f, 559.2f), P(0, 548.8f, 559.2f), P(0, 548.8f, 0), white, center, true);
    quad(P(549.6f, 0, 559.2f), P(0, 0, 559.2f), P(0, 548.8f, 559.2f), P(556.0f, 548.8f, 559.2f), white, center, true);
    quad(P(0, 0, 559.2f), P(0, 0, 0), P(0, 548.8f, 0), P(0, 548.8f, 559.2f), green, center, true);
    quad(P(552.8f, 0, 0), P(549.6f, 0, 559.2f), P(556.0f, 548.8f, 559.2f), P(556.0f, 548.8f, 0), red, center, true);
    const h3 sb[8] = {P(130, 0, 65), P(82, 0, 225), P(240, 0, 272), P(290, 0, 114),
                      P(130, 165, 65), P(82, 165, 225), P(240, 165, 272), P(290, 165, 114)};
    const h3 sbc = (sb[0] + sb[2] + sb[5] + sb[7]) * 0.25f;
    quad(sb[4], sb[5], sb[6], sb[7], box_white, sbc, false);
    quad(sb[0], sb[4], sb[7], sb[3], box_white, sbc, false);
    quad(sb[3], sb[7], sb[6], sb[2], box_white, sbc, false);
    quad(sb[2], sb[6], sb[5], sb[1], box_white, sbc, false);
    quad(sb[1], sb[5], sb[4], sb[0], box_white, sbc, false);
    const h3 tb[8] = {P(423, 0, 247), P(265, 0, 296), P(314, 0, 456), P(472, 0, 406),
                      P(423, 330, 247), P(265, 330, 296), P(314, 330, 456), P(472, 330, 406)};
    const h3 tbc = (tb[0] + tb[2] + tb[5] + tb[7]) * 0.25f;
    quad(tb[4], tb[5], tb[6], tb[7], box_white, tbc, false);
    quad(tb[0], tb[4], tb[7], tb[3], box_white, tbc, false);
    quad(tb[3], tb[7], tb[6], tb[2], box_white, tbc, false);
    quad(tb[2], tb[6], tb[5], tb[1], box_white, tbc, false);
    quad(tb[1], tb[5], tb[4], tb[0], box_white, tbc, false);
}

// ---- scenes/cornell_spheres.h:11-101 ------------------------------------------------
void build_cornell_spheres(HostScene& s) {
    const int white = add_material(s, lambertian(0.78f, 0.78f, 0.78f));
    const int red = add_material(s, lambertian(0.63f, 0.065f, 0.05f));
    const int green = add_material(s, lambertian(0.14f, 0.45f, 0.091f));
    const int light_mat = add_material(s, lambertian(1.0f, 1.0f, 1.0f));
    const int mirror_white = add_material(s, lib_mirror_white());
    const int gold = add_material(s, lib_gold());
    const int silver = add_material(s, lib_silver());
    const float sc = 0.001f;
    auto P = [&](float x, float y, float z) { return H3(x * sc, y * sc, z * sc); };
    const h3 center = P(278.0f, 274.4f, 279.6f);
    // a triangle whose normal is flipped to face the room centre
    auto add_oriented = [&](h3 a, h3 b, h3 c, int mat, h3 em) {
        rpd_triangle t = make_triangle(a, b, c, mat, em);
        const h3 mid = (a + b + c) / 3.0f;
        if (hdot(get3(t.normal), center - mid) < 0.0f) put3(t.normal, -get3(t.normal));
        return add_triangle(s, t);
    };
    auto quad = [&](h3 a, h3 b, h3 c, h3 d, int mat) {
        add_oriented(a, b, c, mat, H3(0, 0, 0));
        add_oriented(a, c, d, mat, H3(0, 0, 0));
    };
    const h3 color = H3(70.0f, 70.0f, 70.0f);
    const h3 lo = P(213.0f, 548.0f, 227.0f);
    const int li = add_light(s, rect_light(lo, P(343.0f, 548.0f, 227.0f) - lo, P(213.0f, 548.0f, 332.0f) - lo, color));
    const h3 l0 = P(213.0f, 548.0f, 227.0f), l1 = P(343.0f, 548.0f, 227.0f), l2 = P(343.0f, 548.0f, 332.0f),
             l3 = P(213.0f, 548.0f, 332.0f);
    link_triangle_to_light(s, add_oriented(l0, l2, l1, light_mat, color), li);
    link_triangle_to_light(s, add_oriented(l0, l3, l2, light_mat, color), li);
    quad(P(552.8f, 0, 0), P(0, 0, 0), P(0, 0, 559.2f), P(549.6f, 0, 559.2f), white);
    quad(P(556.0f, 548.8f, 0), P(556.0f, 548.8f, 559.2f), P(0, 548.8f, 559.2f), P(0, 548.8f, 0), white);
    quad(P(549.6f, 0, 559.2f), P(0, 0, 559.2f), P(0, 548.8f, 559.2f), P(556.0f, 548.8f, 559.2f), white);
    quad(P(0, 0, 559.2f), P(0, 0, 0), P(0, 548.8f, 0), P(0, 548.8f, 559.2f), green);
    quad(P(552.8f, 0, 0), P(549.6f, 0, 559.2f), P(556.0f, 548.8f, 559.2f), P(556.0f, 548.8f, 0), red);
    s.spheres.push_back(sphere(P(150.0f, 90.0f, 150.0f), 0.09f, mirror_white));
    s.spheres.push_back(sphere(P(350.0f, 120.0f, 350.0f), 0.12f, gold));
    s.spheres.push_back(sphere(P(450.0f, 80.0f, 200.0f), 0.08f, silver));
}

// ---- scenes/veach_mis.h:12-108 ------------------------------------------------------
void build_veach_mis(HostScene& s) {
    s.env[0] = s.env[1] = s.env[2] = 0.02f;
    const h3 n2 = H3(1.5f, 1.5f, 1.5f), a = H3(0.9f, 0.9f, 0.9f);
    const int smooth = add_material(s, microfacet(a, 0.001f, n2));
    const int medium1 = add_material(s, microfacet(a, 0.1f, n2));
    const int medium2 = add_material(s, microfacet(a, 0.3f, n2));
    const int rough = add_material(s, microfacet(a, 0.7f, n2));
    const int ground = add_material(s, lambertian(0.3f, 0.3f, 0.3f));
    const int light_mat = add_material(s, lambertian(1.0f, 1.0f, 1.0f));
    add_ground_plane(s, 20.0f, 0.0f, ground);
    const float radius = 0.5f, spacing = 2.0f, start_x = -3.0f;
    const int mats[4] = {smooth, medium1, medium2, rough};
    for (int i = 0; i < 4; ++i) s.spheres.push_back(sphere(H3(start_x + i * spacing, radius, 0.0f), radius, mats[i]));
    const float sizes[4] = {1.2f, 0.6f, 0.3f, 0.15f};
    for (int i = 0; i < 4; ++i) {
        const float x = start_x + i * spacing, size = sizes[i];
        const float intensity = 40.0f / (size * size);
        add_rect_light_quad(s, H3(x - size * 0.5f, 3.0f, -2.0f - size * 0.5f), H3(size, 0.0f, 0.0f),
                            H3(0.0f, 0.0f, size), H3(intensity, intensity, intensity), light_mat);
    }
}

// ---- scenes/material_showcase.h:12-131 ----------------------------------------------
void build_material_showcase(HostScene& s) {
    s.env[0] = s.env[1] = s.env[2] = 0.05f;
    rpd_material copper = lib_copper(), gold = lib_gold(), aluminum = lib_aluminum();
    copper.roughness = 0.05f;
    gold.roughness = 0.15f;
    aluminum.roughness = 0.4f;
    const int red = add_material(s, lambertian(0.8f, 0.1f, 0.1f));
    const int green = add_material(s, lambertian(0.1f, 0.8f, 0.1f));
    const int blue = add_material(s, lambertian(0.1f, 0.1f, 0.8f));
    const int copper_id = add_material(s, copper);
    const int gold_id = add_material(s, gold);
    const int aluminum_id = add_material(s, aluminum);
    const int mirror_id = add_material(s, lib_mirror_white());
    const int jade_id = add_material(s, lib_jade());
    const int silver_id = add_material(s, lib_silver());
    const int ground = add_material(s, lambertian(0.4f, 0.4f, 0.4f));
    const int back_wall = add_material(s, lambertian(0.6f, 0.6f, 0.65f));
    const int light_mat = add_material(s, lambertian(1.0f, 1.0f, 1.0f));
    add_ground_plane(s, 30.0f, 0.0f, ground);
    const h3 w0 = H3(-10.0f, 0.0f, 5.0f), w1 = H3(10.0f, 0.0f, 5.0f), w2 = H3(10.0f, 8.0f, 5.0f), w3 = H3(-10.0f, 8.0f, 5.0f);
    add_triangle(s, make_triangle(w0, w1, w2, back_wall));
    add_triangle(s, make_triangle(w0, w2, w3, back_wall));
    const int grid[3][3] = {{red, green, blue}, {copper_id, gold_id, aluminum_id}, {mirror_id, jade_id, silver_id}};
    for (int row = 0; row < 3; ++row)
        for (int col = 0; col < 3; ++col)
            s.spheres.push_back(sphere(H3(-2.5f + col * 2.5f, 0.6f, -2.0f + row * 2.2f), 0.6f, grid[row][col]));
    add_rect_light_quad(s, H3(-3.0f, 4.0f, -2.0f), H3(1.5f, 0.0f, 0.0f), H3(0.0f, 0.0f, 1.5f), H3(50.0f, 50.0f, 50.0f),
                        light_mat);
    add_rect_light_quad(s, H3(3.0f, 3.5f, -1.0f), H3(0.8f, 0.0f, 0.0f), H3(0.0f, 0.0f, 0.8f), H3(20.0f, 20.0f, 20.0f),
                        light_mat);
    add_light(s, point_light(H3(0.0f, 3.0f, 4.0f), H3(15.0f, 15.0f, 15.0f)));
}

// ---- scenes/caustics.h:12-118 -------------------------------------------------------
void build_caustics(HostScene& s) {
    s.env[0] = s.env[1] = s.env[2] = 0.01f;
    const int mirror_id = add_material(s, lib_mirror_white());
    const int ground = add_material(s, lambertian(0.7f, 0.7f, 0.7f));
    const int wall = add_material(s, lambertian(0.6f, 0.6f, 0.65f));
    const int accent = add_material(s, lambertian(0.2f, 0.3f, 0.5f));
    const int light_mat = add_material(s, lambertian(1.0f, 1.0f, 1.0f));
    add_ground_plane(s, 25.0f, 0.0f, ground);
    auto wall_quad = [&](h3 a, h3 b, h3 c, h3 d) {
        add_triangle(s, make_triangle(a, b, c, wall));
        add_triangle(s, make_triangle(a, c, d, wall));
    };
    wall_quad(H3(-8.0f, 0.0f, 6.0f), H3(8.0f, 0.0f, 6.0f), H3(8.0f, 8.0f, 6.0f), H3(-8.0f, 8.0f, 6.0f));
    wall_quad(H3(-8.0f, 0.0f, -5.0f), H3(-8.0f, 0.0f, 6.0f), H3(-8.0f, 8.0f, 6.0f), H3(-8.0f, 8.0f, -5.0f));
    wall_quad(H3(8.0f, 0.0f, 6.0f), H3(8.0f, 0.0f, -5.0f), H3(8.0f, 8.0f, -5.0f), H3(8.0f, 8.0f, 6.0f));
    wall_quad(H3(-8.0f, 8.0f, -5.0f), H3(8.0f, 8.0f, -5.0f), H3(8.0f, 8.0f, 6.0f), H3(-8.0f, 8.0f, 6.0f));
    s.spheres.push_back(sphere(H3(0.0f, 1.2f, 0.0f), 1.2f, mirror_id));
    s.spheres.push_back(sphere(H3(-2.5f, 0.7f, -1.5f), 0.7f, mirror_id));
    s.spheres.push_back(sphere(H3(2.5f, 0.6f, -1.0f), 0.6f, mirror_id));
    s.spheres.push_back(sphere(H3(-1.5f, 0.35f, 2.0f), 0.35f, accent));
    add_light(s, point_light(H3(-3.0f, 6.0f, -2.0f), H3(200.0f, 200.0f, 200.0f)));
    add_light(s, point_light(H3(3.5f, 5.0f, -1.0f), H3(80.0f, 80.0f, 80.0f)));
    add_rect_light_quad(s, H3(-0.8f, 7.5f, -0.8f), H3(1.6f, 0.0f, 0.0f), H3(0.0f, 0.0f, 1.6f), H3(40.0f, 40.0f, 40.0f),
                        light_mat);
}

// ---- scenes/debug_scene.h:9-57 ------------------------------------------------------
void build_debug(HostScene& s) {
    const int floor_id = add_material(s, lambertian(0.8f, 0.8f, 0.8f));
    const int left = add_material(s, lambertian(0.2f, 0.6f, 0.9f));
    const int right = add_material(s, lambertian(0.9f, 0.3f, 0.3f));
    add_triangle(s, make_triangle(H3(-1.0f, -1.0f, -1.0f), H3(1.0f, -1.0f, -1.0f), H3(1.0f, -1.0f, -3.0f), floor_id));
    add_triangle(s, make_triangle(H3(-1.0f, -1.0f, -1.0f), H3(1.0f, -1.0f, -3.0f), H3(-1.0f, -1.0f, -3.0f), floor_id));
    add_triangle(s, make_triangle(H3(-0.2f, -1.0f, -1.5f), H3(-0.2f, 0.2f, -1.5f), H3(-0.8f, 0.2f, -2.5f), left));
    add_triangle(s, make_triangle(H3(-0.2f, -1.0f, -1.5f), H3(-0.8f, 0.2f, -2.5f), H3(-0.8f, -1.0f, -2.5f), left));
    add_triangle(s, make_triangle(H3(0.5f, -1.0f, -2.0f), H3(0.9f, 0.5f, -2.2f), H3(0.1f, 0.4f, -2.5f), right));
}

// ---- the mesh scenes: scenes/bunny.h:14-148, bunny_dof.h:15-163, dragon.h:14-153 ---
// Their shared room, in the scenes' source frame mapped by (x, y, z) -> (x, z, -y):
// a rect light with two emitting triangles, then back, bottom, left, right and
// top walls of two triangles each.
h3 to_scene(float x, float y, float z) { return H3(x, z, -y); }

struct MeshRoom {
    float wall_albedo[6][3];  // back, bottom, left, right, top, light
    float light;              // emitted radiance (grey)
};

void add_mesh_room(HostScene& s, const MeshRoom& r, int* mat_ids, bool accent_point_light) {
    for (int k = 0; k < 6; ++k) mat_ids[k] = add_material(s, lambertian(r.wall_albedo[k][0], r.wall_albedo[k][1], r.wall_albedo[k][2]));
    const float lx = 0.195f, ly = -0.355f, lz = 0.545f, llx = 0.16f, lly = 0.16f;
    const h3 color = H3(r.light, r.light, r.light);
    const int li = add_light(s, rect_light(to_scene(lx, ly, lz), H3(llx, 0.0f, 0.0f), H3(0.0f, 0.0f, lly), color));
    const int t1 = add_triangle(s, make_triangle(to_scene(lx, ly + lly, lz), to_scene(lx + llx, ly, lz), to_scene(lx, ly, lz),
                                                 mat_ids[5], color));
    link_triangle_to_light(s, t1, li);
    const int t2 = add_triangle(s, make_triangle(to_scene(lx, ly + lly, lz), to_scene(lx + llx, ly + lly, lz),
                                                 to_scene(lx + llx, ly, lz), mat_ids[5], color));
    link_triangle_to_light(s, t2, li);
    if (accent_point_light)  // dragon.h:69-72
        add_light(s, point_light(to_scene(0.278f, -0.05f, 0.24f), H3(0.2f, 0.2f, 0.24f)));
    const float x0 = 0.000000133f, x1 = 0.555999935f, y0 = -0.559199989f, y1 = -0.000000119f;
    const float z0 = 0.000000040f, z1 = 0.548799932f;
    const h3 w[10][3] = {
        {to_scene(x0, y0, z1), to_scene(x1, y0, z0), to_scene(x0, y0, z0)},   // back
        {to_scene(x0, y0, z1), to_scene(x1, y0, z1), to_scene(x1, y0, z0)},
        {to_scene(x0, y0, z0), to_scene(x1, y0, z0), to_scene(x1, y1, z0)},   // bottom
        {to_scene(x0, y0, z0), to_scene(x1, y1, z0), to_scene(x0, y1, z0)},
        {to_scene(x1, y1, z1), to_scene(x1, y1, z0), to_scene(x1, y0, z0)},   // left
        {to_scene(x1, y1, z1), to_scene(x1, y0, z0), to_scene(x1, y0, z1)},
        {to_scene(x0, y0, z0), to_scene(x0, y1, z0), to_scene(x0, y1, z1)},   // right
        {to_scene(x0, y0, z0), to_scene(x0, y1, z1), to_scene(x0, y0, z1)},
        {to_scene(x0, y1, z1), to_scene(x1, y0, z1), to_scene(x0, y0, z1)},   // top
        {to_scene(x0, y1, z1), to_scene(x1, y1, z1), to_scene(x1, y0, z1)},
    };
    for (int k = 0; k < 10; ++k) add_triangle(s, make_triangle(w[k][0], w[k][1], w[k][2], mat_ids[k / 2]));
}

struct MeshFrame { float center_x, center_z, min_y, height; };

MeshFrame mesh_frame(const std::vector<rpd_triangle>& mesh) {
    h3 lo = H3(FLT_MAX, FLT_MAX, FLT_MAX), hi = H3(-FLT_MAX, -FLT_MAX, -FLT_MAX);
    for (const rpd_triangle& t : mesh)
        for (const float* v : {t.v0, t.v1, t.v2}) {
            lo = H3(std::min(lo.x, v[0]), std::min(lo.y, v[1]), std::min(lo.z, v[2]));
            hi = H3(std::max(hi.x, v[0]), std::max(hi.y, v[1]), std::max(hi.z, v[2]));
        }
    return {0.5f * (lo.x + hi.x), 0.5f * (lo.z + hi.z), lo.y, std::max(kEpsSmall, hi.y - lo.y)};
}

// one placed copy: centre on x/z, floor at y = 0, scale, yaw about +y, translate
void add_mesh_instance(HostScene& s, const std::vector<rpd_triangle>& mesh, const MeshFrame& f, float scale, float angle,
                       h3 translate, int material) {
    const float c = std::cos(angle), sn = std::sin(angle);
    for (rpd_triangle t : mesh) {
        h3 v[3];
        const float* src[3] = {t.v0, t.v1, t.v2};
        for (int i = 0; i < 3; ++i) {
            h3 p = H3(src[i][0] - f.center_x, src[i][1] - f.min_y, src[i][2] - f.center_z) * scale;
            p = H3(c * p.x + sn * p.z, p.y, -sn * p.x + c * p.z);
            v[i] = p + translate;
        }
        put3(t.v0, v[0]); put3(t.v1, v[1]); put3(t.v2, v[2]);
        put3(t.normal, hnorm(hcross(v[1] - v[0], v[2] - v[0])));
        t.material_id = material;
        add_triangle(s, t);
    }
}

int build_bunny(HostScene& s, std::string& err) {
    const MeshRoom room = {{{0.874000013f, 0.874000013f, 0.875000000f}, {0.874000013f, 0.874000013f, 0.875000000f},
                            {0.0f, 0.2117f, 0.3765f}, {0.996f, 0.7373f, 0.0667f},
                            {0.874000013f, 0.874000013f, 0.875000000f}, {1.0f, 1.0f, 1.0f}}, 80.0f};
    int ids[6];
    const int gold = 6;  // added right after the six room materials
    std::vector<rpd_triangle> mesh;
    if (int st = load_obj_triangles("resources/models/bunny.obj", gold, mesh, err)) return st;
    add_mesh_room(s, room, ids, false);
    if (add_material(s, lib_gold()) != gold) { err = "material order"; return RPD_ERR_STATE; }
    const MeshFrame f = mesh_frame(mesh);
    add_mesh_instance(s, mesh, f, 0.32f / f.height, float(M_PI), H3(0.28f, 1e-3f, 0.25f), gold);
    return RPD_OK;
}

int build_bunny_dof(HostScene& s, std::string& err) {
    const MeshRoom room = {{{0.874f, 0.874f, 0.875f}, {0.874f, 0.874f, 0.875f}, {0.0f, 0.2117f, 0.3765f},
                            {0.996f, 0.7373f, 0.0667f}, {0.874f, 0.874f, 0.875f}, {1.0f, 1.0f, 1.0f}}, 50.0f};
    std::vector<rpd_triangle> mesh;
    if (int st = load_obj_triangles("resources/models/bunny.obj", 0, mesh, err)) return st;
    int ids[6];
    add_mesh_room(s, room, ids, false);
    // bunny_dof.h:41-48 (the sphere materials are registered but no sphere uses them)
    add_material(s, lambertian(0.75f, 0.2f, 0.2f));
    add_material(s, lambertian(0.2f, 0.75f, 0.2f));
    add_material(s, lambertian(0.2f, 0.35f, 0.9f));
    add_material(s, lib_gold());
    add_material(s, lib_copper());
    const int diffuse = add_material(s, lambertian(0.8f, 0.7f, 0.6f));
    const int gold = add_material(s, lib_gold());
    const int copper = add_material(s, lib_copper());
    const MeshFrame f = mesh_frame(mesh);
    const float face = float(M_PI);
    struct Inst { float x, y, z, height, yaw; int mat; };
    const Inst inst[5] = {{0.18f, 1e-3f, 0.20f, 0.14f, 0.1f, diffuse}, {0.34f, 1e-3f, 0.34f, 0.12f, -0.2f, gold},
                          {0.22f, 1e-3f, 0.46f, 0.15f, 0.25f, copper}, {0.42f, 1e-3f, 0.52f, 0.13f, -0.35f, diffuse},
                          {0.30f, 1e-3f, 0.62f, 0.11f, 0.4f, gold}};
    for (const Inst& in : inst)
        add_mesh_instance(s, mesh, f, 1.2f * in.height / f.height, face + in.yaw, H3(in.x, in.y, in.z), in.mat);
    return RPD_OK;
}

int build_dragon(HostScene& s, std::string& err) {
    const MeshRoom room = {{{0.874000013f, 0.874000013f, 0.875000000f}, {0.874000013f, 0.874000013f, 0.875000000f},
                            {0.05f, 0.15f, 0.25f}, {0.6f, 0.4f, 0.08f},
                            {0.874000013f, 0.874000013f, 0.875000000f}, {1.0f, 1.0f, 1.0f}}, 30.0f};
    const int jade = 6;
    std::vector<rpd_triangle> mesh;
    if (int st = load_obj_triangles("resources/models/dragon.obj", jade, mesh, err)) return st;
    int ids[6];
    add_mesh_room(s, room, ids, true);
    if (add_material(s, lib_jade()) != jade) { err = "material order"; return RPD_ERR_STATE; }
    const MeshFrame f = mesh_frame(mesh);
    add_mesh_instance(s, mesh, f, 0.32f / f.height, float(1.5 * M_PI), H3(0.28f, 1e-3f, 0.32f), jade);
    return RPD_OK;
}

// One vertex reference of an OBJ face: "i", "i/j", "i//k" or "i/j/k", 1-based,
// negative counting back from the latest vertex. Returns false if malformed.
bool obj_index(const std::string& tok, int n_vertices, int& out) {
    const char* p = tok.c_str();
    char* end = nullptr;
    const long v = std::strtol(p, &end, 10);
    if (end == p || (*end != '\0' && *end != '/')) return false;
    const long i = v > 0 ? v - 1 : v < 0 ? n_vertices + v : -1;
    if (i < 0 || i >= n_vertices) return false;
    out = int(i);
    return true;
}

}  // namespace

int load_obj_triangles(const std::string& path, int material_id, std::vector<rpd_triangle>& out, std::string& err) {
    std::ifstream in(path);
    if (!in) {
        err = "cannot open OBJ file '" + path + "'";
        return RPD_ERR_INVALID_ARG;
    }
    std::vector<h3> verts;
    std::string line;
    int line_no = 0;
    while (std::getline(in, line)) {
        ++line_no;
        std::istringstream ls(line);
        std::string key;
        if (!(ls >> key)) continue;
        if (key == "v") {
            float xyz[3];
            std::string tok;
            for (float& c : xyz) {
                if (!(ls >> tok)) { err = path + ":" + std::to_string(line_no) + ": vertex with fewer than 3 coordinates"; return RPD_ERR_INVALID_ARG; }
                c = float(std::strtod(tok.c_str(), nullptr));
            }
            verts.push_back(H3(xyz[0], xyz[1], xyz[2]));
        } else if (key == "f") {
            std::vector<int> idx;
            std::string tok;
            while (ls >> tok) {
                int i = 0;
                if (!obj_index(tok, int(verts.size()), i)) {
                    err = path + ":" + std::to_string(line_no) + ": bad face index '" + tok + "'";
                    return RPD_ERR_INVALID_ARG;
                }
                idx.push_back(i);
            }
            if (idx.size() < 3) { err = path + ":" + std::to_string(line_no) + ": face with fewer than 3 vertices"; return RPD_ERR_INVALID_ARG; }
            for (size_t k = 1; k + 1 < idx.size(); ++k) {
                const h3 a = verts[size_t(idx[0])], b = verts[size_t(idx[k])], c = verts[size_t(idx[k + 1])];
                rpd_triangle t{};
                put3(t.v0, a); put3(t.v1, b); put3(t.v2, c);
                put3(t.normal, hnorm(hcross(b - a, c - a)));  // obj_loader.h:15-17
                t.material_id = material_id;
                t.light_id = -1;
                out.push_back(t);
            }
        }
    }
    return RPD_OK;
}

int build_builtin_scene(const std::string& name, HostScene& s, std::string& err) {
    s = HostScene{};
    if (name == "cornell_box") { build_cornell(s); return RPD_OK; }
    if (name == "debug") { build_debug(s); return RPD_OK; }
    if (name == "cornell_spheres") { build_cornell_spheres(s); return RPD_OK; }
    if (name == "veach_mis") { build_veach_mis(s); return RPD_OK; }
    if (name == "material_showcase") { build_material_showcase(s); return RPD_OK; }
    if (name == "caustics") { build_caustics(s); return RPD_OK; }
    if (name == "bunny") return build_bunny(s, err);
    if (name == "bunny_dof") return build_bunny_dof(s, err);
    if (name == "dragon") return build_dragon(s, err);
    err = "Unknown scene '" + name + "'. Available scenes:\n"
          "  Primitives: cornell_box, cornell_spheres, veach_mis, material_showcase, caustics\n"
          "  Meshes: bunny, bunny_dof, dragon\n"
          "  Debug: debug";
    return RPD_ERR_INVALID_ARG;
}

}  // namespace rpd_host
