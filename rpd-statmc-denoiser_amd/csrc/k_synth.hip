// Synthetic per-sample records for the denoise-only workload, generated on
// the device by a counter-based generator: every value is a pure function of
// (seed, pixel, sample, dimension), so records of any pixel can be produced in
// any order and reproduced exactly.
//
// Image model: square regions of `region` pixels, each a plane with its own
// unit normal, depth (uniform in [1, 20]) and albedo (uniform in [0.1, 0.9]);
// per-sample radiance is lognormal (log-mean mu, log-sigma sigma) tinted by the
// albedo, with a fraction of fireflies scaled up; ten random dimensions
// uniform in [0, 1], three of them (lens x, brdf x, light x) tied to the
// radiance's normal deviate through a Gaussian copula with correlation rho.
#include "kernels.h"

namespace rpd {

// SplitMix64 output function (Steele, Lea, Flood, OOPSLA 2014).
__device__ __forceinline__ uint64_t splitmix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}
// uniform in [0, 1) with 24 random bits
__device__ __forceinline__ float synth_u(uint64_t key, uint32_t dim) {
    return float(splitmix64(key + 0x9e3779b97f4a7c15ull * (uint64_t(dim) + 1)) >> 40) * 0x1p-24f;
}
// standard normal deviate by Box-Muller from two uniform dimensions
__device__ __forceinline__ float synth_normal(uint64_t key, uint32_t dim) {
    const float u1 = synth_u(key, dim) + 0x1p-25f, u2 = synth_u(key, dim + 1);
    return sqrtf(-2.0f * logf(u1)) * cospif(2.0f * u2);
}
// standard normal CDF
__device__ __forceinline__ float phi(float x) { return 0.5f * erfcf(-x * 0.70710678118654752f); }

__global__ void k_synth(SynthParams sp, Records rec) {
    const int64_t slot = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    const int64_t n = int64_t(sp.W) * sp.H * sp.spp;
    if (slot >= n) return;
    const int p = int(slot / sp.spp), s = int(slot % sp.spp);
    const int x = p % sp.W, y = p / sp.W;
    const uint64_t base = uint64_t(sp.seed) << 32;
    // region plane
    const int rx = x / sp.region, ry = y / sp.region;
    const uint64_t rkey = splitmix64(base ^ (0x5ull << 60) ^ (uint64_t(uint32_t(ry)) << 20) ^ uint64_t(uint32_t(rx)));
    v3 nr = mk3(2.0f * synth_u(rkey, 0) - 1.0f, 2.0f * synth_u(rkey, 1) - 1.0f, 1.0f);
    nr = normalized3(nr);
    const float depth = 1.0f + 19.0f * synth_u(rkey, 2);
    const v3 alb = mk3(0.1f + 0.8f * synth_u(rkey, 3), 0.1f + 0.8f * synth_u(rkey, 4), 0.1f + 0.8f * synth_u(rkey, 5));
    // sample
    const uint64_t key = splitmix64(base ^ (uint64_t(uint32_t(p)) << 24) ^ uint64_t(uint32_t(s)));
    const float z = synth_normal(key, 0);
    float lum = expf(sp.mu + sp.sigma * z);
    if (synth_u(key, 2) < sp.ff_frac) lum *= sp.ff_scale;
    const float amean = (alb.x + alb.y + alb.z) * (1.0f / 3.0f);
    const v3 rgb = (lum / amean) * alb;
    const float c = sqrtf(fmaxf(0.0f, 1.0f - sp.rho * sp.rho));
    float d[10];
    for (int k = 0; k < 10; ++k) d[k] = synth_u(key, 3 + k);
    // correlated: lens x (0), brdf x (2), light x (4)
    d[0] = fminf(phi(sp.rho * z + c * synth_normal(key, 13)), 0x1.fffffep-1f);
    d[2] = fminf(phi(sp.rho * z + c * synth_normal(key, 15)), 0x1.fffffep-1f);
    d[4] = fminf(phi(sp.rho * z + c * synth_normal(key, 17)), 0x1.fffffep-1f);
    const float pu_x = synth_u(key, 19), pu_y = synth_u(key, 20);
    const int light = min(int(d[8] * 4.0f), 3);
    const bool survived = d[9] < 0.8f;
    uint32_t flags = REC_HIT | REC_PIXEL | REC_LENS | REC_LSEL | REC_BRDF | REC_LIGHT | REC_ENV | REC_RR;
    if (survived) flags |= REC_RR_SURVIVED;
    if (finite3(rgb)) flags |= REC_FINITE;
    flags |= uint32_t(light) << REC_LIGHT_SHIFT;
    rec.r0[slot] = make_float4(rgb.x, rgb.y, rgb.z, __uint_as_float(flags));
    rec.r1[slot] = make_float4(nr.x, nr.y, nr.z, depth);
    rec.r2[slot] = make_float4(alb.x, alb.y, alb.z, 0.0f);
    rec.r3[slot] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    rec.r4[slot] = make_float4(pu_x, pu_y, d[0], d[1]);
    rec.r5[slot] = make_float4(d[2], d[3], d[4], d[5]);
    rec.r6[slot] = make_float2(d[6], d[7]);
}

void launch_synth(const SynthParams& sp, const Records& rec, hipStream_t st) {
    const int64_t n = int64_t(sp.W) * sp.H * sp.spp;
    if (n <= 0) return;
    hipLaunchKernelGGL(k_synth, dim3(unsigned((n + 255) / 256)), dim3(256), 0, st, sp, rec);
}

}  // namespace rpd
