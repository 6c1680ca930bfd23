// ORACLE — test infrastructure only. This directory is a CPU restatement of
// the reference renderer's hot path (path trace -> RPF grid -> StatMC filter ->
// variance smoothing -> adaptive allocation). Only tests/, the smoke check and
// the CPU-baseline leg of bench.py may load it; the product never does.
//
// Every expression below fixes the evaluation order and float/double
// promotions that the reference's compiler (GCC, x86-64 baseline, no FMA)
// gives the reference's Eigen fixed-size expressions:
//   * 3-term reductions (dot, squaredNorm) unroll as x0 + (x1 + x2)
//     (Eigen redux_novec_unroller splits Length=3 as 1 + 2);
//   * normalized()/normalize() divide each component by sqrt(squaredNorm),
//     leaving a zero vector unchanged;
//   * `float op M_PI` expressions are evaluated in double and rounded once;
//   * Vec3f op double scalars convert the scalar to float first.
#pragma once
#include <cmath>
#include <cstdint>
#include <algorithm>

#define ORC_EPS 1e-4f        // path_tracer/core/common.h:5
#define ORC_EPS_ANYHIT 5e-4f // path_tracer/core/common.h:6
#define ORC_EPS_SMALL 1e-8f  // shared/global.h:4

namespace orc {

struct F2 { float x = 0.0f, y = 0.0f; };

struct F3 {
    float x = 0.0f, y = 0.0f, z = 0.0f;
    F3() = default;
    constexpr F3(float a, float b, float c) : x(a), y(b), z(c) {}
    float operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
    float& operator[](int i) { return i == 0 ? x : (i == 1 ? y : z); }
};

inline F3 operator+(F3 a, F3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline F3 operator-(F3 a, F3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline F3 operator-(F3 a) { return {-a.x, -a.y, -a.z}; }
inline F3 operator*(F3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
inline F3 operator*(float s, F3 a) { return {s * a.x, s * a.y, s * a.z}; }
inline F3 operator/(F3 a, float s) { return {a.x / s, a.y / s, a.z / s}; }
inline F3& operator+=(F3& a, F3 b) { a = a + b; return a; }
inline F3 mul(F3 a, F3 b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
inline F3 divv(F3 a, F3 b) { return {a.x / b.x, a.y / b.y, a.z / b.z}; }
inline float dot(F3 a, F3 b) { return a.x * b.x + (a.y * b.y + a.z * b.z); }
inline float sqnorm(F3 a) { return a.x * a.x + (a.y * a.y + a.z * a.z); }
inline float norm(F3 a) { return std::sqrt(sqnorm(a)); }
inline F3 normalized(F3 a) {
    const float z = sqnorm(a);
    if (z > 0.0f) { const float s = std::sqrt(z); return {a.x / s, a.y / s, a.z / s}; }
    return a;
}
inline F3 cross(F3 a, F3 b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
inline float max_coeff(F3 a) { return std::max(a.x, std::max(a.y, a.z)); }
inline bool all_finite(F3 a) { return std::isfinite(a.x) && std::isfinite(a.y) && std::isfinite(a.z); }
inline bool is_zero(F3 a, float prec) {
    return std::fabs(a.x) <= prec && std::fabs(a.y) <= prec && std::fabs(a.z) <= prec;
}
inline F3 fmax3(F3 a, float m) { return {std::max(a.x, m), std::max(a.y, m), std::max(a.z, m)}; }

// shared/global.h:45-48
inline float lerpf(float a, float b, float t) { return a + (b - a) * t; }
inline F3 lerp3(F3 a, F3 b, float t) { return a + (b - a) * t; }

// core/common.h:37-39
inline float luminance(F3 c) { return 0.2126f * c.x + 0.7152f * c.y + 0.0722f * c.z; }

// core/common.h:14-35 (Pixar revised orthonormal basis)
inline void onb(F3 n, F3& bx, F3& by) {
    const float sign = std::copysign(1.0f, n.z);
    const float a = -1.0f / (sign + n.z);
    const float b = n.x * n.y * a;
    bx = {1.0f + sign * n.x * n.x * a, sign * b, -sign * n.x};
    by = {b, sign + n.y * n.y * a, -n.y};
}
inline F3 local_to_world(F3 v, F3 n) {
    F3 bx, by; onb(n, bx, by);
    return v.x * bx + v.y * by + v.z * n;
}
inline F3 world_to_local(F3 v, F3 n) {
    F3 bx, by; onb(n, bx, by);
    return {dot(v, bx), dot(v, by), dot(v, n)};
}
// core/common.h:33-35
inline F3 offset_origin(F3 p, F3 n) { return p + ORC_EPS * n; }

// path_tracer/main.cpp:270-286
inline uint32_t hash_u32(uint32_t x) {
    x ^= x >> 16; x *= 0x7feb352du; x ^= x >> 15; x *= 0x846ca68bu; x ^= x >> 16;
    return x;
}
inline float hash_to_unit_float(uint32_t x) {
    return float(hash_u32(x) & 0x00ffffffu) / float(0x01000000u);
}
inline uint32_t pixel_sample_seed(uint32_t base, int pixel_index, int sample_offset) {
    return hash_u32(base ^ hash_u32(uint32_t(pixel_index) + 0x9e3779b9u) ^
                    hash_u32(uint32_t(sample_offset) + 0x85ebca6bu));
}
// main.cpp:266-268
inline float relative_depth_delta(float a, float b) {
    return std::fabs(a - b) / std::max(std::max(std::fabs(a), std::fabs(b)), ORC_EPS_SMALL);
}

}  // namespace orc
