// Procedural benchmark scene (host code): an icosphere subdivided `level`
// times (20 * 4^level flat triangles) whose vertices are displaced along the
// radius by fractal value noise, spheres with GGX microfacet materials
// scattered in a shell around it, rectangular area lights facing it and a
// constant environment. Every value derives from the seed through integer
// hashing, so the same configuration always produces the same bits.
#include "../../include/rpd_engine.h"

#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace {

struct P3 { float x, y, z; };
P3 operator+(P3 a, P3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
P3 operator-(P3 a, P3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
P3 operator*(float s, P3 a) { return {s * a.x, s * a.y, s * a.z}; }
P3 cross(P3 a, P3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
float dot(P3 a, P3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
P3 normalize(P3 a) {
    const float l = std::sqrt(dot(a, a));
    return {a.x / l, a.y / l, a.z / l};
}

uint32_t mix32(uint32_t x) {  // integer finaliser (the seed hash of the reference uses the same form)
    x ^= x >> 16; x *= 0x7feb352du; x ^= x >> 15; x *= 0x846ca68bu; x ^= x >> 16;
    return x;
}
float unit(uint32_t h) { return float(mix32(h) >> 8) * (1.0f / 16777216.0f); }

// value noise on the integer lattice, trilinear with smoothstep weights, in [0, 1)
float lattice(int x, int y, int z, uint32_t seed) {
    return unit(uint32_t(x) * 0x8da6b343u ^ uint32_t(y) * 0xd8163841u ^ uint32_t(z) * 0xcb1ab31fu ^ seed);
}
float value_noise(P3 p, uint32_t seed) {
    const float fx = std::floor(p.x), fy = std::floor(p.y), fz = std::floor(p.z);
    const int ix = int(fx), iy = int(fy), iz = int(fz);
    const float tx = p.x - fx, ty = p.y - fy, tz = p.z - fz;
    const float ux = tx * tx * (3.0f - 2.0f * tx), uy = ty * ty * (3.0f - 2.0f * ty), uz = tz * tz * (3.0f - 2.0f * tz);
    float c[2][2][2];
    for (int a = 0; a < 2; ++a)
        for (int b = 0; b < 2; ++b)
            for (int d = 0; d < 2; ++d) c[a][b][d] = lattice(ix + a, iy + b, iz + d, seed);
    auto lerp = [](float a, float b, float t) { return a + (b - a) * t; };
    const float x00 = lerp(c[0][0][0], c[1][0][0], ux), x10 = lerp(c[0][1][0], c[1][1][0], ux);
    const float x01 = lerp(c[0][0][1], c[1][0][1], ux), x11 = lerp(c[0][1][1], c[1][1][1], ux);
    return lerp(lerp(x00, x10, uy), lerp(x01, x11, uy), uz);
}
// octaves of noise at doubling frequency and halving weight, in [-1, 1]
float fbm(P3 p, int octaves, uint32_t seed) {
    float sum = 0.0f, wsum = 0.0f, w = 1.0f, f = 1.0f;
    for (int o = 0; o < octaves; ++o) {
        sum += w * (2.0f * value_noise(f * p, seed + 0x9e3779b9u * uint32_t(o + 1)) - 1.0f);
        wsum += w;
        w *= 0.5f;
        f *= 2.0f;
    }
    return wsum > 0.0f ? sum / wsum : 0.0f;
}

void put(float* d, P3 v) { d[0] = v.x; d[1] = v.y; d[2] = v.z; }

struct Gen {
    std::vector<rpd_triangle> tris;
    std::vector<rpd_sphere> spheres;
    std::vector<rpd_material> mats;
    std::vector<rpd_light> lights;
};

int build(const rpd_procedural_config& c, Gen& g) {
    if (c.subdivision_level < 0 || c.subdivision_level > 9 || c.n_spheres < 0 || c.noise_octaves < 0)
        return RPD_ERR_INVALID_ARG;
    // materials: 0 icosphere, 1 light quads, 2.. spheres
    rpd_material ico{};
    ico.kind = RPD_MAT_LAMBERTIAN;
    ico.albedo[0] = 0.72f; ico.albedo[1] = 0.62f; ico.albedo[2] = 0.5f;
    g.mats.push_back(ico);
    rpd_material dark{};
    dark.kind = RPD_MAT_LAMBERTIAN;
    g.mats.push_back(dark);
    // icosahedron (12 vertices, 20 faces wound counter-clockwise seen from outside)
    const float t = (1.0f + std::sqrt(5.0f)) * 0.5f;
    const P3 v[12] = {{-1, t, 0}, {1, t, 0}, {-1, -t, 0}, {1, -t, 0}, {0, -1, t}, {0, 1, t},
                      {0, -1, -t}, {0, 1, -t}, {t, 0, -1}, {t, 0, 1}, {-t, 0, -1}, {-t, 0, 1}};
    const int f[20][3] = {{0, 11, 5}, {0, 5, 1}, {0, 1, 7}, {0, 7, 10}, {0, 10, 11}, {1, 5, 9}, {5, 11, 4},
                          {11, 10, 2}, {10, 7, 6}, {7, 1, 8}, {3, 9, 4}, {3, 4, 2}, {3, 2, 6}, {3, 6, 8},
                          {3, 8, 9}, {4, 9, 5}, {2, 4, 11}, {6, 2, 10}, {8, 6, 7}, {9, 8, 1}};
    std::vector<P3> cur, nxt;
    cur.reserve(60);
    for (const auto& fc : f)
        for (int k = 0; k < 3; ++k) cur.push_back(normalize(v[fc[k]]));
    for (int l = 0; l < c.subdivision_level; ++l) {
        nxt.clear();
        nxt.reserve(cur.size() * 4);
        for (size_t i = 0; i < cur.size(); i += 3) {
            const P3 a = cur[i], b = cur[i + 1], d = cur[i + 2];
            // a + b is commutative in floating point, so a shared edge gets the same midpoint in both faces
            const P3 ab = normalize(a + b), bd = normalize(b + d), da = normalize(d + a);
            const P3 out[12] = {a, ab, da, ab, b, bd, da, bd, d, ab, bd, da};
            nxt.insert(nxt.end(), out, out + 12);
        }
        cur.swap(nxt);
    }
    const uint32_t nseed = mix32(c.seed ^ 0x6e6f6973u);
    auto displace = [&](P3 p) {
        const float r = 1.0f + c.displacement * fbm(c.noise_frequency * p, c.noise_octaves, nseed);
        return r * p;
    };
    g.tris.reserve(cur.size() / 3 + 8);
    for (size_t i = 0; i < cur.size(); i += 3) {
        const P3 a = displace(cur[i]), b = displace(cur[i + 1]), d = displace(cur[i + 2]);
        P3 n = cross(b - a, d - a);
        const float l2 = dot(n, n);
        n = l2 > 0.0f ? normalize(n) : P3{0.0f, 1.0f, 0.0f};
        rpd_triangle tr{};
        put(tr.v0, a); put(tr.v1, b); put(tr.v2, d); put(tr.normal, n);
        tr.material_id = 0;
        tr.light_id = -1;
        g.tris.push_back(tr);
    }
    // area lights: quads on a ring above the object, facing its centre
    const int n_lights = 4;
    for (int k = 0; k < n_lights; ++k) {
        const float ang = 6.2831853f * (float(k) + 0.125f) / float(n_lights);
        const P3 centre{4.0f * std::cos(ang), 3.5f, 4.0f * std::sin(ang)};
        const P3 n = normalize(P3{0.0f, 0.0f, 0.0f} - centre);
        P3 u = normalize(cross(n, P3{0.0f, 1.0f, 0.0f}));
        P3 w = cross(u, n);
        if (dot(cross(u, w), n) < 0.0f) { const P3 s = u; u = w; w = s; }
        const float size = 1.4f;
        const P3 width = size * u, length = size * w;
        const P3 corner = centre - 0.5f * (width + length);
        rpd_light L{};
        L.kind = RPD_LIGHT_RECT;
        put(L.p0, corner); put(L.p1, width); put(L.p2, length); put(L.normal, n);
        const float e = c.light_emission * (0.8f + 0.4f * unit(c.seed * 31u + uint32_t(k)));
        L.emission[0] = e; L.emission[1] = e * 0.95f; L.emission[2] = e * 0.85f;
        g.lights.push_back(L);
        const P3 q[4] = {corner, corner + width, corner + width + length, corner + length};
        const int tri_idx[2][3] = {{0, 1, 2}, {0, 2, 3}};
        for (const auto& ti : tri_idx) {
            rpd_triangle tr{};
            put(tr.v0, q[ti[0]]); put(tr.v1, q[ti[1]]); put(tr.v2, q[ti[2]]); put(tr.normal, n);
            std::memcpy(tr.emission, L.emission, sizeof(L.emission));
            tr.material_id = 1;
            tr.light_id = k;
            g.tris.push_back(tr);
        }
    }
    // spheres in a shell around the object, each with its own GGX material
    for (int k = 0; k < c.n_spheres; ++k) {
        const uint32_t s = c.seed * 0x9e3779b1u + uint32_t(k) * 0x85ebca77u;
        const float z = 1.0f - 2.0f * unit(s + 1u), ph = 6.2831853f * unit(s + 2u);
        const float rr = std::sqrt(std::fmax(0.0f, 1.0f - z * z));
        const float dist = 1.6f + 1.8f * unit(s + 3u);
        rpd_sphere sp{};
        sp.center[0] = dist * rr * std::cos(ph); sp.center[1] = dist * z; sp.center[2] = dist * rr * std::sin(ph);
        sp.radius = 0.05f + 0.17f * unit(s + 4u);
        sp.material_id = int(g.mats.size());
        sp.light_id = -1;
        g.spheres.push_back(sp);
        rpd_material m{};
        m.kind = RPD_MAT_MICROFACET;
        m.distribution = RPD_MICROFACET_GGX;
        m.roughness = 0.05f + 0.75f * unit(s + 5u);
        for (int ch = 0; ch < 3; ++ch) {
            m.albedo[ch] = 0.1f + 0.8f * unit(s + 6u + uint32_t(ch));
            m.n1[ch] = 1.0f;
            m.n2[ch] = 1.5f;
        }
        g.mats.push_back(m);
    }
    return RPD_OK;
}

}  // namespace

extern "C" {

void rpd_procedural_config_defaults(rpd_procedural_config* c) {
    if (!c) return;
    std::memset(c, 0, sizeof(*c));
    c->seed = 7;
    c->subdivision_level = 8;
    c->n_spheres = 256;
    c->displacement = 0.18f;
    c->noise_frequency = 1.6f;
    c->noise_octaves = 4;
    c->environment[0] = 0.05f; c->environment[1] = 0.06f; c->environment[2] = 0.08f;
    c->light_emission = 14.0f;
}

int rpd_procedural_scene_arrays(const rpd_procedural_config* c, rpd_triangle* tris, int* n_tris, rpd_sphere* spheres,
                                int* n_spheres, rpd_material* mats, int* n_mats, rpd_light* lights, int* n_lights,
                                float env[3]) {
    if (!c || !n_tris || !n_spheres || !n_mats || !n_lights) return RPD_ERR_INVALID_ARG;
    Gen g;
    const int st = build(*c, g);
    if (st) return st;
    const bool fill = tris && spheres && mats && lights;
    if (fill && (*n_tris < int(g.tris.size()) || *n_spheres < int(g.spheres.size()) || *n_mats < int(g.mats.size()) ||
                 *n_lights < int(g.lights.size())))
        return RPD_ERR_INVALID_ARG;
    *n_tris = int(g.tris.size()); *n_spheres = int(g.spheres.size());
    *n_mats = int(g.mats.size()); *n_lights = int(g.lights.size());
    if (fill) {
        std::memcpy(tris, g.tris.data(), g.tris.size() * sizeof(rpd_triangle));
        std::memcpy(spheres, g.spheres.data(), g.spheres.size() * sizeof(rpd_sphere));
        std::memcpy(mats, g.mats.data(), g.mats.size() * sizeof(rpd_material));
        std::memcpy(lights, g.lights.data(), g.lights.size() * sizeof(rpd_light));
    }
    if (env) std::memcpy(env, c->environment, 3 * sizeof(float));
    return RPD_OK;
}

int rpd_scene_procedural(rpd_context* ctx, const rpd_procedural_config* c, rpd_scene** out) {
    if (!c || !out) return RPD_ERR_INVALID_ARG;
    Gen g;
    const int st = build(*c, g);
    if (st) return st;
    return rpd_scene_create(ctx, g.tris.data(), int(g.tris.size()), g.spheres.data(), int(g.spheres.size()),
                            g.mats.data(), int(g.mats.size()), g.lights.data(), int(g.lights.size()), c->environment,
                            out);
}

}  // extern "C"
