// Device-side math, sampler and scene helpers for gfx950.
//
// All kernels are compiled with -ffp-contract=off: the reference renderer is
// built for the x86-64 baseline ISA (no FMA), so every a*b+c there rounds
// twice and parity needs the same here. Fixed evaluation orders below follow
// the reference's Eigen fixed-size expressions (3-term sums as x0 + (x1 + x2),
// per-component division in normalize) and its float/double promotions.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string.h>

#define RPD_EPS 1e-4f         // path_tracer/core/common.h:5
#define RPD_EPS_ANYHIT 5e-4f  // path_tracer/core/common.h:6
#define RPD_EPS_SMALL 1e-8f   // shared/global.h:4
#define RPD_PI 3.14159265358979323846

namespace rpd {

// Plain float helpers used by both the kernels and host-side checks of the
// intersection code (same flags on both sides: no contraction, IEEE sqrt/div).
#define RPD_HD __host__ __device__ __forceinline__

struct v2 { float x, y; };
struct v3 { float x, y, z; };

RPD_HD v3 mk3(float a, float b, float c) { v3 r; r.x = a; r.y = b; r.z = c; return r; }
RPD_HD v3 operator+(v3 a, v3 b) { return mk3(a.x + b.x, a.y + b.y, a.z + b.z); }
RPD_HD v3 operator-(v3 a, v3 b) { return mk3(a.x - b.x, a.y - b.y, a.z - b.z); }
RPD_HD v3 operator-(v3 a) { return mk3(-a.x, -a.y, -a.z); }
RPD_HD v3 operator*(v3 a, float s) { return mk3(a.x * s, a.y * s, a.z * s); }
RPD_HD v3 operator*(float s, v3 a) { return mk3(s * a.x, s * a.y, s * a.z); }
RPD_HD v3 operator/(v3 a, float s) { return mk3(a.x / s, a.y / s, a.z / s); }
RPD_HD v3 vmul(v3 a, v3 b) { return mk3(a.x * b.x, a.y * b.y, a.z * b.z); }
RPD_HD v3 vdiv(v3 a, v3 b) { return mk3(a.x / b.x, a.y / b.y, a.z / b.z); }
RPD_HD float dot3(v3 a, v3 b) { return a.x * b.x + (a.y * b.y + a.z * b.z); }
RPD_HD float sqnorm3(v3 a) { return a.x * a.x + (a.y * a.y + a.z * a.z); }
RPD_HD float norm3(v3 a) { return sqrtf(sqnorm3(a)); }
RPD_HD v3 normalized3(v3 a) {
    const float z = sqnorm3(a);
    if (z > 0.0f) { const float s = sqrtf(z); return mk3(a.x / s, a.y / s, a.z / s); }
    return a;
}
RPD_HD v3 cross3(v3 a, v3 b) {
    return mk3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
RPD_HD float maxc3(v3 a) { return fmaxf(a.x, fmaxf(a.y, a.z)); }
RPD_HD bool finite3(v3 a) { return isfinite(a.x) && isfinite(a.y) && isfinite(a.z); }
RPD_HD float comp3(v3 a, int i) { return i == 0 ? a.x : (i == 1 ? a.y : a.z); }
RPD_HD float clampf(float v, float lo, float hi) { return fminf(fmaxf(v, lo), hi); }
RPD_HD float lerpf_(float a, float b, float t) { return a + (b - a) * t; }
// Correctly rounded expf: exp in double, rounded once to float. The
// reference's std::exp(float) is glibc's expf, which agrees with the correctly
// rounded value except on roughly one argument in 10^4.
RPD_HD float expf_cr(float x) { return float(exp(double(x))); }
RPD_HD float luminance3(v3 c) { return 0.2126f * c.x + 0.7152f * c.y + 0.0722f * c.z; }

// core/common.h:14-35: Pixar revised orthonormal basis
RPD_HD void onb(v3 n, v3& bx, v3& by) {
    const float sign = copysignf(1.0f, n.z);
    const float a = -1.0f / (sign + n.z);
    const float b = n.x * n.y * a;
    bx = mk3(1.0f + sign * n.x * n.x * a, sign * b, -sign * n.x);
    by = mk3(b, sign + n.y * n.y * a, -n.y);
}
RPD_HD v3 local_to_world(v3 v, v3 n) {
    v3 bx, by; onb(n, bx, by);
    return v.x * bx + v.y * by + v.z * n;
}
RPD_HD v3 world_to_local(v3 v, v3 n) {
    v3 bx, by; onb(n, bx, by);
    return mk3(dot3(v, bx), dot3(v, by), dot3(v, n));
}
RPD_HD v3 offset_origin(v3 p, v3 n) { return p + RPD_EPS * n; }

// path_tracer/main.cpp:266-286
__host__ __device__ __forceinline__ uint32_t hash_u32(uint32_t x) {
    x ^= x >> 16; x *= 0x7feb352du; x ^= x >> 15; x *= 0x846ca68bu; x ^= x >> 16;
    return x;
}
__host__ __device__ __forceinline__ float hash_to_unit_float(uint32_t x) {
    return float(hash_u32(x) & 0x00ffffffu) / float(0x01000000u);
}
__host__ __device__ __forceinline__ uint32_t pixel_sample_seed(uint32_t base, int idx, int n) {
    return hash_u32(base ^ hash_u32(uint32_t(idx) + 0x9e3779b9u) ^ hash_u32(uint32_t(n) + 0x85ebca6bu));
}
__device__ __forceinline__ float relative_depth_delta(float a, float b) {
    return fabsf(a - b) / fmaxf(fmaxf(fabsf(a), fabsf(b)), RPD_EPS_SMALL);
}

// ---- MT19937 tape -----------------------------------------------------------
// The reference seeds a std::mt19937 once per pixel per pass and draws
// std::uniform_real_distribution<float> values from it (core/sampler.h:7-16).
// A tape holds one pixel's 624-word generator state right after a twist; draws
// temper one word each, and a twist regenerates the tape in place.
constexpr int MT_N = 624;
constexpr int MT_M = 397;
constexpr uint32_t MT_UPPER = 0x80000000u;
constexpr uint32_t MT_LOWER = 0x7fffffffu;
constexpr uint32_t MT_MATRIX = 0x9908b0dfu;

__device__ __forceinline__ uint32_t mt_mix(uint32_t hi_word, uint32_t lo_word, uint32_t far_word) {
    const uint32_t y = (hi_word & MT_UPPER) | (lo_word & MT_LOWER);
    return far_word ^ (y >> 1) ^ ((y & 1u) ? MT_MATRIX : 0u);
}
__device__ __forceinline__ uint32_t mt_temper(uint32_t y) {
    y ^= (y >> 11);
    y ^= (y << 7) & 0x9d2c5680u;
    y ^= (y << 15) & 0xefc60000u;
    y ^= (y >> 18);
    return y;
}
// libstdc++ generate_canonical<float,24> with a 32-bit engine: float(u)/2^32,
// a result of 1 is replaced by the largest float below 1.
__device__ __forceinline__ float mt_to_unit(uint32_t u) {
    float r = float(u) / 4294967296.0f;
    return r >= 1.0f ? 0x1.fffffep-1f : r;
}
// Tapes are stored in groups of 64: word w of tape a lives at
//   (a / 64) * 756 * 64 + (w / 4) * 256 + (a % 64) * 4 + (w % 4)
// so 64 consecutive tapes write one contiguous KiB per 4 words when they are
// generated together, and one tape still reads 4 consecutive words as a uint4.
// Words 0-623: the state after the first twist (or its first `prefix` words).
// Words 624-625 (MT_TAPE_CURSOR) of a prefix: the seeding recurrence's x[prefix]
// and x[prefix + 397], from which the trace extends the prefix. Words 628-755
// of a whole tape: the first MT_S2 words of the state after the second twist,
// so that a pixel drawing a little past word 623 needs no twist of its own.
constexpr int MT_GROUP = 64;
constexpr int MT_TAPE_CURSOR = MT_N;
constexpr int MT_S2 = 128;
constexpr int MT_S2_AT = MT_N + 4;
constexpr int MT_TAPE_WORDS = MT_S2_AT + MT_S2;
__host__ __device__ __forceinline__ size_t tape_base(size_t a) {
    return (a / MT_GROUP) * size_t(MT_TAPE_WORDS * MT_GROUP) + (a % MT_GROUP) * 4;
}
__device__ __forceinline__ int tape_word(int w) { return (w >> 2) * (4 * MT_GROUP) + (w & 3); }
// Tape word of draw position pos: 0-623 the current state, 624-751 the next
// state's prefix (tape words 628-755).
__device__ __forceinline__ int tape_pos_word(int pos) { return tape_word(pos < MT_N ? pos : pos + (MT_S2_AT - MT_N)); }

__host__ __device__ __forceinline__ uint32_t mt_seed_step(uint32_t prev, uint32_t k) {
    return 1812433253u * (prev ^ (prev >> 30)) + k;
}

// The words of a pixel's tape written before the trace: room for `count`
// samples of per_sample draws plus margin words, in whole 4-word chunks, at
// most the full 624. A lane that draws past them writes the rest itself
// (mt_tape_extend); the words are the same either way.
__host__ __device__ __forceinline__ int tape_prefix_words(int count, int per_sample, int margin) {
    const long long w = (long long)per_sample * count + margin;
    return w >= MT_N ? MT_N : int((w + 3) / 4 * 4);
}
// Draw positions readable from a tape with that prefix: a whole tape also
// holds the next state's first MT_S2 words.
__host__ __device__ __forceinline__ int tape_draw_limit(int prefix) { return prefix >= MT_N ? MT_N + MT_S2 : prefix; }

// Writes words [from, 624) of the tape after its first twist, words [0, from)
// and the seeding cursors x[from], x[from + 397] (MT_TAPE_CURSOR) being there
// already (from a multiple of 4): the seeded state is walked on from the two
// cursors (x[k] and, while k < 227, x[k+397]) and words from 227 on mix in the
// tape word 227 back. Rare (a pixel whose samples drew more than its prefix),
// so written for size, not speed.
__device__ __noinline__ void mt_tape_extend(uint32_t* __restrict__ s, int from);

// The standard twist over all 624 words, in place.
__device__ inline void mt_twist(uint32_t* __restrict__ s) {
    for (int k = 0; k < MT_N - MT_M; ++k)
        s[tape_word(k)] = mt_mix(s[tape_word(k)], s[tape_word(k + 1)], s[tape_word(k + MT_M)]);
    for (int k = MT_N - MT_M; k < MT_N - 1; ++k)
        s[tape_word(k)] = mt_mix(s[tape_word(k)], s[tape_word(k + 1)], s[tape_word(k + MT_M - MT_N)]);
    s[tape_word(MT_N - 1)] = mt_mix(s[tape_word(MT_N - 1)], s[tape_word(0)], s[tape_word(MT_M - 1)]);
}

__device__ __noinline__ void mt_tape_extend(uint32_t* __restrict__ s, int from) {
    constexpr int kBack = MT_N - MT_M;  // 227
    uint32_t lo = s[tape_word(MT_TAPE_CURSOR)];        // x[from]
    uint32_t far = s[tape_word(MT_TAPE_CURSOR + 1)];   // x[from + 397], used while k < 227
    for (int k = from; k < MT_N - 1; ++k) {
        const uint32_t x1 = mt_seed_step(lo, uint32_t(k + 1));
        const uint32_t back = k < kBack ? far : s[tape_word(k - kBack)];
        s[tape_word(k)] = mt_mix(lo, x1, back);
        lo = x1;
        if (k + 1 < kBack) far = mt_seed_step(far, uint32_t(k + 1 + MT_M));
    }
    s[tape_word(MT_N - 1)] = mt_mix(lo, s[tape_word(0)], s[tape_word(MT_M - 1)]);
}

// The rare refill of a draw cursor at the end of the draws its tape holds
// (pos == limit, out of line):
//  * pos < 624, past the prefix k_tape wrote: the lane writes the rest of the
//    state, limit becomes 624;
//  * pos == 624 of a tape without the next state's prefix (extended, or
//    twisted here before): the tape is twisted in place and the draws go on
//    from its word 0;
//  * pos == 624 + MT_S2, past the next state's prefix: the tape (still the
//    first state) is twisted in place and the draws go on from its word
//    MT_S2; the prefix is gone, limit becomes 624.
struct TapeRefill { uint4 cache; int pos; int limit; };
__device__ __noinline__ TapeRefill mt_refill_slow(uint32_t* __restrict__ tape, int pos, int limit) {
    TapeRefill r;
    r.pos = pos;
    r.limit = MT_N;
    if (pos < MT_N) {
        mt_tape_extend(tape, pos);
        r.cache = *reinterpret_cast<const uint4*>(tape + tape_word(pos));
    } else {
        mt_twist(tape);
        r.pos = pos - MT_N;
        r.cache = *reinterpret_cast<const uint4*>(tape + tape_word(r.pos));
    }
    return r;
}

// Per-lane draw cursor over a tape with a 4-word register cache.
struct Draws {
    uint32_t* tape;  // tape_base of this pixel's tape
    int pos;         // next draw position in [0, limit] (tape_pos_word)
    int limit;       // draw positions the tape holds (a multiple of 4; tape_draw_limit)
    uint4 cache;     // words [pos & ~3, pos | 3]
    __device__ __forceinline__ void reset(uint32_t* t, int prefix) { tape = t; pos = 0; limit = tape_draw_limit(prefix); }
    __device__ __forceinline__ float next() {
        if ((pos & 3) == 0) {
            if (pos >= limit) {   // pos == limit: extend the tape or twist it (rare)
                const TapeRefill r = mt_refill_slow(tape, pos, limit);
                cache = r.cache;
                pos = r.pos;
                limit = r.limit;
            } else {
                cache = *reinterpret_cast<const uint4*>(tape + tape_pos_word(pos));
            }
        }
        const int k = pos & 3;
        const uint32_t w = k == 0 ? cache.x : (k == 1 ? cache.y : (k == 2 ? cache.z : cache.w));
        ++pos;
        return mt_to_unit(mt_temper(w));
    }
    // Vec2f(next1d(), next1d()) as GCC evaluates it: x is the second draw.
    __device__ __forceinline__ v2 next2() {
        const float first = next();
        const float second = next();
        v2 r; r.x = second; r.y = first; return r;
    }
};

// core/sampler.h:22-35
__device__ __forceinline__ v2 sample_disk(v2 u) {
    const float ox = 2.0f * u.x - 1.0f, oy = 2.0f * u.y - 1.0f;
    v2 r; r.x = 0.0f; r.y = 0.0f;
    if (ox == 0.0f && oy == 0.0f) return r;
    float theta, rad;
    if (fabsf(ox) > fabsf(oy)) { rad = ox; theta = float((RPD_PI / 4.0) * double(oy / ox)); }
    else { rad = oy; theta = float((RPD_PI / 2.0) - (RPD_PI / 4.0) * double(ox / oy)); }
    r.x = rad * cosf(theta); r.y = rad * sinf(theta);
    return r;
}
// core/sampler.h:37-43
__device__ __forceinline__ v3 sample_sphere(v2 u) {
    const float z = 1.0f - 2.0f * u.x;
    const float r = sqrtf(fmaxf(0.0f, 1.0f - z * z));
    const float phi = float(2.0 * RPD_PI * double(u.y));
    return mk3(r * cosf(phi), r * sinf(phi), z);
}

// ---- scene ------------------------------------------------------------------
enum : int { MAT_LAMBERT = 0, MAT_MICROFACET = 1, MAT_MIRROR = 2 };
enum : int { LIGHT_RECT = 0, LIGHT_AREA = 1, LIGHT_POINT = 2, LIGHT_DIRECTIONAL = 3 };

struct DevMaterial {  // 48 B
    int kind, beckmann;
    float roughness, pad;
    float albedo[3], n1[3], n2[3], pad2[3];
};
struct DevLight {  // 80 B
    int kind;
    float area, power, pad;
    float p0[3], p1[3], p2[3], normal[3], emission[3], pad2;
};
// Triangles as 16 floats: v0, v1, v2, normal, then material id, light id
// (-1 if not an emitter) as integer bits.
struct DevTriangle { float v0[3], v1[3], v2[3], n[3]; int material, light, emitter, pad; };
struct DevSphere { float c[3], r; int material, light, emitter, pad; };

// 4-wide BVH node: bounds of the four children as structure of arrays (one
// 16-byte load per axis and side), then the links. count > 0: a leaf of that
// many primitive slots starting at child; count == 0: inner node child;
// count < 0: no child.
struct Bvh4Node {
    float lo_x[4], hi_x[4], lo_y[4], hi_y[4], lo_z[4], hi_z[4];
    int child[4];
    int count[4];
};
constexpr int BVH_STACK = 48;        // traversal stack entries per lane
constexpr int BVH_MAX_DEPTH4 = 16;   // the builder keeps 4-wide depth within this
constexpr int BVH_SLOT_FLOAT4 = 9;   // per primitive slot: triangle (3 axis orders x 3 float4) or sphere (1 float4)

struct DevScene {
    const DevTriangle* tris; int n_tris;
    const float4* tri_perm;  // n_tris * 3 entries of 3 float4, see tri_perm_entry
    const float4* tri_attr;  // per triangle: what shading reads, see tri_attr_pack
    const DevSphere* spheres; int n_spheres;
    const DevMaterial* materials;
    const DevLight* lights; int n_lights;
    int n_materials;
    const float* light_cdf;  // inclusive cumulative power
    float env[3];
    int env_on;
    int lambert_only;        // every material is Lambertian: the trace kernel's specialised variant applies
};

// What shading reads of a hit triangle, in one float4: the geometric normal,
// then the material id (low 16 bits), the light id + 1 (next 15 bits; 0: not
// linked to a light) and the emitter flag (top bit) as integer bits.
constexpr int TRI_ATTR_MAX_MATERIALS = 1 << 16;
constexpr int TRI_ATTR_MAX_LIGHTS = (1 << 15) - 1;
struct TriShade { v3 n; int material, light; bool emitter; };
RPD_HD float4 tri_attr_pack(const DevTriangle& t) {
    const uint32_t bits = uint32_t(t.material) | (uint32_t(t.light + 1) << 16) | (t.emitter ? 0x80000000u : 0u);
    float4 a;
    a.x = t.n[0]; a.y = t.n[1]; a.z = t.n[2];
    memcpy(&a.w, &bits, 4);
    return a;
}
RPD_HD TriShade tri_attr_unpack(float4 a) {
    TriShade t;
    t.n = mk3(a.x, a.y, a.z);
    uint32_t bits;
    memcpy(&bits, &a.w, 4);
    t.material = int(bits & 0xffffu);
    t.light = int((bits >> 16) & 0x7fffu) - 1;
    t.emitter = (bits >> 31) != 0;
    return t;
}

// BVH over a scene's triangles and spheres, for scenes too large for the flat
// loop. A separate block so that the flat-loop kernel's arguments keep their
// layout whether or not a scene has one.
struct DevBvh {
    const Bvh4Node* nodes;   // null: the scene has no BVH
    const float4* slots;     // BVH_SLOT_FLOAT4 float4 per primitive slot, in leaf order
    const int* slot_prim;    // triangle index, or sphere index | 0x80000000
};

RPD_HD v3 ld3(const float* p) { return mk3(p[0], p[1], p[2]); }

// Watertight ray/triangle test (Woop, Benthin, Wald, JCGT 2013), as in
// path_tracer/core/triangle.h:35-104.
//
// The test works in a ray space whose z axis kz is the ray's dominant axis.
// The reference also swaps kx and ky when d[kz] < 0 to keep the winding; that
// swap exchanges the sheared x and y coordinates, which negates U, V, W, det
// and T exactly (IEEE a - b is exactly -(b - a)), so t and the hit decision do
// not depend on it. Only the three cyclic orders (kx, ky, kz) =
// (kz+1, kz+2, kz) mod 3 are used, and the triangle vertices are stored
// pre-permuted for each of them: one table lookup replaces nine per-lane
// component selects.
struct RayPre {
    int kz;
    float Sx, Sy, Sz;
    float okx, oky, okz;
};
RPD_HD RayPre ray_pre(v3 o, v3 d) {
    RayPre r;
    const float ax = fabsf(d.x), ay = fabsf(d.y), az = fabsf(d.z);
    int kz = 0;
    if (ay > ax && ay > az) kz = 1;
    if (az > ax && az > ay) kz = 2;
    const int kx = kz == 2 ? 0 : kz + 1, ky = kx == 2 ? 0 : kx + 1;
    r.kz = kz;
    const float dkz = comp3(d, kz);
    r.Sx = comp3(d, kx) / dkz;
    r.Sy = comp3(d, ky) / dkz;
    r.Sz = 1.0f / dkz;
    r.okx = comp3(o, kx); r.oky = comp3(o, ky); r.okz = comp3(o, kz);
    return r;
}

// Table entry (tri, kz): v0, v1, v2 with components in the order
// (kz+1, kz+2, kz) mod 3, packed as 9 floats in 3 float4 (3 floats padding).
__host__ __device__ __forceinline__ int tri_perm_entry(int tri, int kz) { return (tri * 3 + kz) * 3; }

RPD_HD bool tri_hit(float4 q0, float4 q1, float4 q2, const RayPre& r, float t_min, float t_max,
                                        float& t_out) {
    const float Akx = q0.x - r.okx, Aky = q0.y - r.oky, Akz = q0.z - r.okz;
    const float Bkx = q0.w - r.okx, Bky = q1.x - r.oky, Bkz = q1.y - r.okz;
    const float Ckx = q1.z - r.okx, Cky = q1.w - r.oky, Ckz = q2.x - r.okz;
    const float Ax = Akx - r.Sx * Akz, Ay = Aky - r.Sy * Akz;
    const float Bx = Bkx - r.Sx * Bkz, By = Bky - r.Sy * Bkz;
    const float Cx = Ckx - r.Sx * Ckz, Cy = Cky - r.Sy * Ckz;
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
    const float T = U * (r.Sz * Akz) + V * (r.Sz * Bkz) + W * (r.Sz * Ckz);
    const float t = T * (1.0f / det);
    if (t < t_min || t > t_max) return false;
    t_out = t;
    return true;
}

// path_tracer/core/sphere.h:19-54
RPD_HD bool sphere_hit(const DevSphere& s, v3 o, v3 d, float t_min, float t_max, float& t_out, v3& n_out) {
    const v3 oc = o - ld3(s.c);
    const float a = sqnorm3(d);
    const float hb = dot3(oc, d);
    const float c = sqnorm3(oc) - s.r * s.r;
    const float disc = hb * hb - a * c;
    if (disc < 0.0f) return false;
    const float sd = sqrtf(disc);
    float t = (-hb - sd) / a;
    if (t < t_min || t > t_max) {
        t = (-hb + sd) / a;
        if (t < t_min || t > t_max) return false;
    }
    n_out = normalized3((o + t * d) - ld3(s.c));
    t_out = t;
    return true;
}

struct HitRec {
    int hit;       // 0 miss
    int sphere;    // 1 if a sphere
    int index;
    float t;
    v3 sn;         // sphere normal
};

// One ray query over all triangles, then all spheres.
//
// Closest-hit mode (t_max = inf) is closest_hit_combined
// (path_tracer/core/ray_tracer.h:304-325): the nearest triangle, the nearest
// sphere root above 1e-8, and the triangle wins only if it is strictly nearer;
// ties keep the lower index. Occlusion mode is any_hit_combined
// (ray_tracer.h:65-75, 282-296, 326-344) with t_max already reduced by 5e-4:
// a triangle blocks when 0 <= t < t_max, a sphere when its root is in
// [1e-8, t_max]. In occlusion mode `hit` reports whether the ray is blocked.
// `tp` is the pre-permuted triangle table (tri_perm_entry), in LDS or HBM.
RPD_HD HitRec trace_query(const DevScene& S, const float4* tp, v3 o, v3 d, float t_max,
                                              bool occlusion) {
    HitRec h; h.hit = 0; h.sphere = 0; h.index = -1; h.t = 0.0f; h.sn = mk3(0, 0, 0);
    const RayPre rp = ray_pre(o, d);
    float best = t_max;
    int bi = -1;
    const float4* e = tp + tri_perm_entry(0, rp.kz);
#pragma unroll 4
    for (int i = 0; i < S.n_tris; ++i, e += 9) {
        float t;
        if (tri_hit(e[0], e[1], e[2], rp, 0.0f, best, t) && t < best) { best = t; bi = i; }
    }
    float sbest = t_max;
    int si = -1;
    v3 sn = mk3(0, 0, 0);
    for (int i = 0; i < S.n_spheres; ++i) {
        float t; v3 n;
        if (sphere_hit(S.spheres[i], o, d, RPD_EPS_SMALL, sbest, t, n) && (t < sbest || occlusion)) {
            sbest = t; si = i; sn = n;
        }
    }
    if (bi >= 0 && (si < 0 || best < sbest)) { h.hit = 1; h.index = bi; h.t = best; }
    else if (si >= 0) { h.hit = 1; h.sphere = 1; h.index = si; h.t = sbest; h.sn = sn; }
    return h;
}

// Traversal stack of node indices, entry k at base[k * stride] (LDS or a
// host array); pushes past BVH_STACK entries are dropped, which the builder's
// depth limit rules out.
struct StridedStack {
    int* base;
    int stride;
    int sp;
    RPD_HD bool empty() const { return sp == 0; }
    RPD_HD void push(int v) { if (sp < BVH_STACK) base[(sp++) * stride] = v; }
    RPD_HD int pop() { return base[(--sp) * stride]; }
};

// The same query through the BVH. Nodes are visited near child first. Triangle
// ties go to the lower index and sphere ties to the lower index, as in the
// index-ordered loop, so the result does not depend on the visiting order; and
// every culling test is widened by 2^-20 so that rounding in the slab test
// never drops a primitive the exact tests accept.
// The query in resumable form: init, then step (one node visit with the
// primitives of its leaf children) until step reports the end, then result.
// A kernel can interleave other work between the steps of a lane's walk.
// A step first tests the four child boxes, then the primitives of every leaf
// child hit in one loop, so the lanes of a wave test primitives together
// whichever child slots their leaves sit in (a C2 job's trace: 786 -> 704 ms
// against one primitive loop per child slot).
struct BvhWalk {
    RayPre rp;
    float ix, iy, iz;
    float best, sbest;
    int bi, si;
    v3 sn;
    int node;
    bool occlusion;
    bool any;  // occlusion: a hit ended the walk

    RPD_HD void init(v3 o, v3 d, float t_max, bool occl) {
        rp = ray_pre(o, d);
        ix = 1.0f / (d.x != 0.0f ? d.x : copysignf(1e-30f, d.x));
        iy = 1.0f / (d.y != 0.0f ? d.y : copysignf(1e-30f, d.y));
        iz = 1.0f / (d.z != 0.0f ? d.z : copysignf(1e-30f, d.z));
        best = t_max; sbest = t_max;
        bi = -1; si = -1;
        sn = mk3(0, 0, 0);
        node = 0;
        occlusion = occl;
        any = false;
    }

    // One node visit for the ray (o, d) given to init; true once the walk has ended.
    template <class Stack>
    RPD_HD bool step(const DevBvh& B, v3 o, v3 d, Stack& stk) {
        const float widen = 1.0f + 0x1p-20f;
        const Bvh4Node& N = B.nodes[node];
        const float cull = fminf(best, sbest) * widen;
        int cn[4];
        float ct[4];
        int nc = 0;
        // leaf children hit: their first slots and counts (3 bits each), tested
        // after the box tests in one loop over all their primitives
        int ls0 = 0, ls1 = 0, ls2 = 0, ls3 = 0, lcnt = 0, nl = 0, ntot = 0;
        for (int k = 0; k < 4; ++k) {
            const int cnt = N.count[k];
            if (cnt < 0) continue;
            const float ax = (N.lo_x[k] - o.x) * ix, bx = (N.hi_x[k] - o.x) * ix;
            const float ay = (N.lo_y[k] - o.y) * iy, by = (N.hi_y[k] - o.y) * iy;
            const float az = (N.lo_z[k] - o.z) * iz, bz = (N.hi_z[k] - o.z) * iz;
            const float tn = fmaxf(fmaxf(fminf(ax, bx), fminf(ay, by)), fmaxf(fminf(az, bz), 0.0f));
            const float tf = fminf(fminf(fmaxf(ax, bx), fmaxf(ay, by)), fmaxf(az, bz));
            if (!(tn <= tf * widen) || !(tn <= cull)) continue;
            if (cnt == 0) { cn[nc] = N.child[k]; ct[nc] = tn; ++nc; continue; }
            const int c0 = N.child[k];
            if (nl == 0) ls0 = c0; else if (nl == 1) ls1 = c0; else if (nl == 2) ls2 = c0; else ls3 = c0;
            lcnt |= cnt << (3 * nl);
            ++nl;
            ntot += cnt;
        }
        {
            int leaf = 0, q = 0;
            for (int i = 0; i < ntot; ++i) {
                while (q >= ((lcnt >> (3 * leaf)) & 7)) { q = 0; ++leaf; }
                const int slot = (leaf == 0 ? ls0 : leaf == 1 ? ls1 : leaf == 2 ? ls2 : ls3) + q;
                ++q;
                const int prim = B.slot_prim[slot];
                const float4* e = B.slots + size_t(slot) * BVH_SLOT_FLOAT4;
                if (prim >= 0) {
                    const float4* pe = e + rp.kz * 3;
                    float t;
                    if (tri_hit(pe[0], pe[1], pe[2], rp, 0.0f, best, t) && (t < best || (bi >= 0 && prim < bi))) {
                        best = t; bi = prim;
                    }
                } else {
                    const int sidx = prim & 0x7fffffff;
                    DevSphere sph;
                    sph.c[0] = e[0].x; sph.c[1] = e[0].y; sph.c[2] = e[0].z; sph.r = e[0].w;
                    float t; v3 n;
                    if (sphere_hit(sph, o, d, RPD_EPS_SMALL, sbest, t, n) &&
                        (occlusion || t < sbest || (si >= 0 && sidx < si))) {
                        sbest = t; si = sidx; sn = n;
                    }
                }
                if (occlusion && (bi >= 0 || si >= 0)) { any = true; return true; }
            }
        }
        // nearest inner child next, the others on the stack farthest first
        for (int a = 1; a < nc; ++a)
            for (int b = a; b > 0 && ct[b] < ct[b - 1]; --b) {
                const float tt = ct[b]; ct[b] = ct[b - 1]; ct[b - 1] = tt;
                const int tc = cn[b]; cn[b] = cn[b - 1]; cn[b - 1] = tc;
            }
        for (int a = nc - 1; a >= 1; --a) stk.push(cn[a]);
        if (nc > 0) { node = cn[0]; return false; }
        if (stk.empty()) return true;
        node = stk.pop();
        return false;
    }

    RPD_HD HitRec result() const {
        HitRec h; h.hit = 0; h.sphere = 0; h.index = -1; h.t = 0.0f; h.sn = mk3(0, 0, 0);
        if (any) { h.hit = 1; return h; }
        if (bi >= 0 && (si < 0 || best < sbest)) { h.hit = 1; h.index = bi; h.t = best; }
        else if (si >= 0) { h.hit = 1; h.sphere = 1; h.index = si; h.t = sbest; h.sn = sn; }
        return h;
    }
};

template <class Stack>
RPD_HD HitRec bvh_query_t(const DevBvh& B, v3 o, v3 d, float t_max, bool occlusion, Stack& stk) {
    BvhWalk w;
    w.init(o, d, t_max, occlusion);
    while (!w.step(B, o, d, stk)) {}
    return w.result();
}
RPD_HD HitRec bvh_query(const DevBvh& B, v3 o, v3 d, float t_max, bool occlusion, int* stack, int stride) {
    StridedStack stk{stack, stride, 0};
    return bvh_query_t(B, o, d, t_max, occlusion, stk);
}

// ---- materials: path_tracer/core/material.h --------------------------------
__device__ __forceinline__ v3 mf_fresnel(const DevMaterial& m, float c) {
    c = clampf(c, 0.0f, 1.0f);
    const v3 n1 = ld3(m.n1), n2 = ld3(m.n2);
    const v3 dn = n1 - n2, sn = n1 + n2;
    const v3 r0 = vdiv(vmul(dn, dn), vmul(sn, sn));
    const float f = powf(1.0f - c, 5.0f);
    return r0 + (mk3(1.0f, 1.0f, 1.0f) - r0) * f;
}
__device__ __forceinline__ float mf_D(const DevMaterial& m, v3 h) {
    const v3 mm = normalized3(h);
    const float c = clampf(mm.z, RPD_EPS, 1.0f);
    const float s2 = fmaxf(0.0f, 1.0f - c * c);
    const float t2 = s2 / (c * c);
    const float a = fmaxf(m.roughness, RPD_EPS);
    if (m.beckmann) {
        const float rs = a * a;
        return float(double(expf_cr(-t2 / rs)) / (RPD_PI * double(rs) * double(powf(c, 4.0f))));
    }
    const float den = c * c * (a * a - 1.0f) + 1.0f;
    return float(double(a * a) / (RPD_PI * double(den) * double(den)));
}
__device__ __forceinline__ float mf_smith(const DevMaterial& m, v3 v) {
    const float c = clampf(v.z, RPD_EPS, 1.0f);
    const float s2 = fmaxf(0.0f, 1.0f - c * c);
    if (s2 <= RPD_EPS_SMALL) return 1.0f;
    const float tn = sqrtf(s2) / c;
    const float a = fmaxf(m.roughness, RPD_EPS);
    if (m.beckmann) {
        const float x = 1.0f / (a * tn);
        if (x >= 1.6f) return 1.0f;
        return (3.535f * x + 2.181f * x * x) / (1.0f + 2.276f * x + 2.577f * x * x);
    }
    const float lambda = (-1.0f + sqrtf(1.0f + a * a * tn * tn)) * 0.5f;
    return 1.0f / (1.0f + lambda);
}
__device__ __forceinline__ float mf_pdf(const DevMaterial& m, v3 wo_w, v3 wi_w, v3 n) {
    const v3 wo = world_to_local(wo_w, n), wi = world_to_local(wi_w, n);
    if (wo.z <= RPD_EPS_SMALL || wi.z <= RPD_EPS_SMALL) return 0.0f;
    const v3 h = normalized3(wo + wi);
    if (!finite3(h)) return 0.0f;
    const float ndoth = fmaxf(0.0f, h.z);
    const float wodoth = fmaxf(0.0f, dot3(wo, h));
    if (ndoth <= RPD_EPS_SMALL || wodoth <= RPD_EPS_SMALL) return 0.0f;
    return mf_D(m, h) * ndoth / (4.0f * wodoth);
}
__device__ __forceinline__ v3 reflect3(v3 d, v3 n) { return d - (2.0f * dot3(d, n)) * n; }

// The BRDF functions take the material kind as an argument: a kernel that
// knows every material of its scene is Lambertian passes MAT_LAMBERT, and the
// other kinds' code drops out of it.
__device__ inline v3 brdf_eval(int kind, const DevMaterial& m, v3 wo_w, v3 wi_w, v3 n) {
    if (kind == MAT_LAMBERT) return ld3(m.albedo) / float(RPD_PI);
    if (kind == MAT_MICROFACET) {
        const v3 wo = normalized3(world_to_local(wo_w, n)), wi = normalized3(world_to_local(wi_w, n));
        if (wo.z <= RPD_EPS_SMALL || wi.z <= RPD_EPS_SMALL) return mk3(0, 0, 0);
        const v3 h = normalized3(wo + wi);
        if (!finite3(h)) return mk3(0, 0, 0);
        const float co = clampf(wo.z, 0.0f, 1.0f), ci = clampf(wi.z, 0.0f, 1.0f);
        const float G = mf_smith(m, wo) * mf_smith(m, wi);
        const v3 fr = ((mf_fresnel(m, fabsf(dot3(wi, h))) * G) * mf_D(m, h)) / (4.0f * ci * co);
        return vmul(fr, ld3(m.albedo));
    }
    const v3 nn = normalized3(n);
    if (dot3(wo_w, nn) <= RPD_EPS_SMALL || dot3(wi_w, nn) <= RPD_EPS_SMALL) return mk3(0, 0, 0);
    const v3 e = normalized3(reflect3(-normalized3(wo_w), nn));
    const v3 in = normalized3(wi_w);
    return (sqnorm3(e - in) <= RPD_EPS_SMALL) ? ld3(m.albedo) : mk3(0, 0, 0);
}
__device__ inline float brdf_pdf(int kind, const DevMaterial& m, v3 wo_w, v3 wi_w, v3 n) {
    if (kind == MAT_LAMBERT) return float(double(fmaxf(0.0f, dot3(wi_w, n))) / RPD_PI);
    if (kind == MAT_MICROFACET) return mf_pdf(m, wo_w, wi_w, n);
    const v3 nn = normalized3(n);
    if (dot3(wo_w, nn) <= RPD_EPS_SMALL || dot3(wi_w, nn) <= RPD_EPS_SMALL) return 0.0f;
    const v3 e = normalized3(reflect3(-normalized3(wo_w), nn));
    const v3 in = normalized3(wi_w);
    return (sqnorm3(e - in) <= RPD_EPS_SMALL) ? 1.0f : 0.0f;
}
__device__ inline void brdf_sample(int kind, const DevMaterial& m, v3 wo_w, v3 n, v2 u, v3& wi, float& pdf) {
    wi = mk3(0, 0, 0); pdf = 0.0f;
    if (kind == MAT_LAMBERT) {
        const float r = sqrtf(u.x);
        const float theta = float(2.0 * RPD_PI * double(u.y));
        const v3 l = mk3(r * cosf(theta), r * sinf(theta), sqrtf(fmaxf(0.0f, 1.0f - u.x)));
        wi = local_to_world(l, n);
        pdf = float(double(l.z) / RPD_PI);
        return;
    }
    if (kind == MAT_MICROFACET) {
        const v3 wo = normalized3(world_to_local(wo_w, n));
        if (wo.z <= RPD_EPS_SMALL) return;
        const float u1 = clampf(u.x, RPD_EPS_SMALL, 1.0f - RPD_EPS_SMALL);
        const float u2 = clampf(u.y, RPD_EPS_SMALL, 1.0f - RPD_EPS_SMALL);
        const float a = fmaxf(m.roughness, RPD_EPS);
        v3 wh;
        if (m.beckmann) {
            const float r = a * sqrtf(-logf(1.0f - u1));
            const float phi = float(2.0 * RPD_PI * double(u2));
            wh = normalized3(mk3(-(r * cosf(phi)), -(r * sinf(phi)), 1.0f));
        } else {
            const float phi = float(2.0 * RPD_PI * double(u1));
            const float ct = sqrtf((1.0f - u2) / (1.0f + (a * a - 1.0f) * u2));
            const float st = sqrtf(fmaxf(0.0f, 1.0f - ct * ct));
            wh = normalized3(mk3(st * cosf(phi), st * sinf(phi), ct));
        }
        if (wh.z <= RPD_EPS_SMALL) wh.z = RPD_EPS;
        const v3 wl = normalized3(-wo + (2.0f * dot3(wo, wh)) * wh);
        if (wl.z <= RPD_EPS_SMALL) return;
        wi = local_to_world(wl, n);
        pdf = mf_pdf(m, wo_w, wi, n);
        return;
    }
    const v3 nn = normalized3(n), wo = normalized3(wo_w);
    if (dot3(wo, nn) <= RPD_EPS_SMALL) return;
    const v3 r = normalized3(reflect3(-wo, nn));
    if (!finite3(r)) return;
    wi = r; pdf = 1.0f;
}
__device__ inline v3 brdf_sample_weight(int kind, const DevMaterial& m, v3 wo, v3 wi, v3 n, float pdf) {
    if (pdf <= RPD_EPS_SMALL) return mk3(0, 0, 0);
    if (kind == MAT_MIRROR) return ld3(m.albedo);
    return brdf_eval(kind, m, wo, wi, n) * (fmaxf(0.0f, dot3(n, wi)) / pdf);
}

// ---- lights: path_tracer/core/light.h, core/scene.h --------------------------
__device__ __forceinline__ bool light_is_delta(const DevLight& l) { return l.kind == LIGHT_POINT || l.kind == LIGHT_DIRECTIONAL; }

__device__ inline void light_sample(const DevLight& l, v3 p, v2 u, v3& wi, float& dist) {
    wi = mk3(0, 0, 0); dist = 0.0f;
    if (l.kind == LIGHT_DIRECTIONAL) { wi = ld3(l.p0); dist = __builtin_inff(); return; }
    v3 target;
    if (l.kind == LIGHT_RECT) {
        target = ld3(l.p0) + u.x * ld3(l.p1) + u.y * ld3(l.p2);
    } else if (l.kind == LIGHT_AREA) {
        const float su = sqrtf(fmaxf(0.0f, u.x));
        const float b0 = 1.0f - su, b1 = u.y * su, b2 = 1.0f - b0 - b1;
        target = b0 * ld3(l.p0) + b1 * ld3(l.p1) + b2 * ld3(l.p2);
    } else {
        target = ld3(l.p0);
    }
    v3 d = target - p;
    const float sq = sqnorm3(d);
    if (sq <= RPD_EPS_SMALL) return;
    d = normalized3(d);
    if (l.kind == LIGHT_RECT || l.kind == LIGHT_AREA) {
        if (dot3(ld3(l.normal), -d) <= RPD_EPS_SMALL) return;
    }
    wi = d;
    dist = sqrtf(sq);
}
__device__ inline v3 eval_light(const DevLight& l, v3 dir, float distance) {
    if (l.kind == LIGHT_RECT || l.kind == LIGHT_AREA)
        return (dot3(dir, ld3(l.normal)) > 0.0f) ? ld3(l.emission) : mk3(0, 0, 0);
    if (l.kind == LIGHT_POINT) return ld3(l.emission) / fmaxf(distance * distance, RPD_EPS_SMALL);
    return ld3(l.emission);
}
__device__ __forceinline__ float light_select_pdf(const DevScene& S, int idx) {
    const float total = S.n_lights > 0 ? S.light_cdf[S.n_lights - 1] : 0.0f;
    if (total <= RPD_EPS_SMALL) return 0.0f;
    return S.lights[idx].power / total;
}
__device__ __forceinline__ float light_pdf(const DevScene& S, int idx, float sel, v3 dir, float dist) {
    const DevLight& l = S.lights[idx];
    float pdf = 1.0f;
    if (l.kind == LIGHT_RECT || l.kind == LIGHT_AREA) {
        const float c = dot3(ld3(l.normal), -dir);
        pdf = (c <= RPD_EPS_SMALL) ? 0.0f : (dist * dist) / (c * l.area);
    }
    return sel * pdf;
}
// lower_bound over the inclusive power CDF (core/scene.h:54-63)
__device__ __forceinline__ void sample_light(const DevScene& S, float u, int& idx, float& pdf) {
    if (S.n_lights == 0) { idx = -1; pdf = 0.0f; return; }
    const float target = u * S.light_cdf[S.n_lights - 1];
    int lo = 0, hi = S.n_lights;
    while (lo < hi) { const int mid = (lo + hi) >> 1; if (S.light_cdf[mid] < target) lo = mid + 1; else hi = mid; }
    if (lo >= S.n_lights) lo = S.n_lights - 1;
    idx = lo;
    pdf = light_select_pdf(S, idx);
}

}  // namespace rpd
