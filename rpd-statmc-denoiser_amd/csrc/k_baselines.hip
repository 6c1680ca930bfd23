// The reference's comparison denoisers on the device:
//   * joint bilateral (denoising/joint_bilateral.cpp:15-128): Gaussian,
//     colour bilateral and joint bilateral outputs from one sweep over the
//     (2r+1)^2 window, r = max(1, ceil(3 sigma_p)), coordinates clamped;
//   * edge-avoiding a-trous wavelet (denoising/atrous_wavelet.cpp:22-183):
//     passes of the 5x5 B3-spline kernel at step 2^iteration.
// Float arithmetic in the reference's order, so outputs differ from it only
// through the float exp/acos of the two math libraries.
//
// The bilateral kernel stages its 16x16 tile and an r-pixel halo in LDS
// (colour, decoded and normalised normal, albedo: 36 B per pixel), so each
// window is read from LDS and every normal is normalised once instead of once
// per window it falls in. Radii beyond BL_LDS_RADIUS read global memory.
// The a-trous passes read neighbours at strides up to 2^(iterations-1)
// through L2, with the normals normalised once and the pass's luminance
// computed once per pixel.
#include "kernels.h"

namespace rpd {

namespace {

constexpr int BL_TILE = 16;
constexpr int BL_LDS_RADIUS = 12;   // (16 + 24)^2 pixels x 36 B = 57.6 KB of LDS

__device__ __forceinline__ void normalize_in_place(float& x, float& y, float& z) {
    const float l2 = x * x + y * y + z * z;
    if (l2 > RPD_EPS_SMALL) {
        const float inv = 1.0f / sqrtf(l2);
        x *= inv; y *= inv; z *= inv;
    }
}

__device__ __forceinline__ void decoded_normal(const float* n, int64_t p, float& x, float& y, float& z) {
    x = n[3 * p] * 2.0f - 1.0f;
    y = n[3 * p + 1] * 2.0f - 1.0f;
    z = n[3 * p + 2] * 2.0f - 1.0f;
    normalize_in_place(x, y, z);
}

// Pixel data of one window sample: colour, unit normal, albedo.
struct BlPix { float r, g, b, nx, ny, nz, ar, ag, ab; };

template <bool kLds>
__global__ void __launch_bounds__(BL_TILE * BL_TILE) k_joint_bilateral(BilateralParams bp) {
    extern __shared__ float s_bl[];
    const int W = bp.width, H = bp.height, r = bp.radius;
    const int tx0 = int(blockIdx.x) * BL_TILE, ty0 = int(blockIdx.y) * BL_TILE;
    const int span = BL_TILE + 2 * r;
    if (kLds) {
        // element (sy, sx) holds the pixel at the clamped coordinates of (ty0 - r + sy, tx0 - r + sx)
        for (int k = threadIdx.x; k < span * span; k += BL_TILE * BL_TILE) {
            const int gy = min(max(ty0 - r + k / span, 0), H - 1), gx = min(max(tx0 - r + k % span, 0), W - 1);
            const int64_t p = int64_t(gy) * W + gx;
            float nx, ny, nz;
            decoded_normal(bp.normal, p, nx, ny, nz);
            float* e = s_bl + 9 * k;
            e[0] = bp.raw[3 * p]; e[1] = bp.raw[3 * p + 1]; e[2] = bp.raw[3 * p + 2];
            e[3] = nx; e[4] = ny; e[5] = nz;
            e[6] = bp.albedo[3 * p]; e[7] = bp.albedo[3 * p + 1]; e[8] = bp.albedo[3 * p + 2];
        }
        __syncthreads();
    }
    const int j = tx0 + int(threadIdx.x) % BL_TILE, i = ty0 + int(threadIdx.x) / BL_TILE;
    if (i >= H || j >= W) return;
    auto fetch = [&](int c, int col) -> BlPix {
        BlPix q;
        if (kLds) {
            const float* e = s_bl + 9 * ((c - ty0 + r) * span + (col - tx0 + r));
            q.r = e[0]; q.g = e[1]; q.b = e[2]; q.nx = e[3]; q.ny = e[4]; q.nz = e[5];
            q.ar = e[6]; q.ag = e[7]; q.ab = e[8];
        } else {
            const int64_t p = int64_t(c) * W + col;
            q.r = bp.raw[3 * p]; q.g = bp.raw[3 * p + 1]; q.b = bp.raw[3 * p + 2];
            decoded_normal(bp.normal, p, q.nx, q.ny, q.nz);
            q.ar = bp.albedo[3 * p]; q.ag = bp.albedo[3 * p + 1]; q.ab = bp.albedo[3 * p + 2];
        }
        return q;
    };
    const BlPix ctr = fetch(i, j);
    const float sp2 = 2.0f * bp.sigma_p * bp.sigma_p, sc2 = 2.0f * bp.sigma_c * bp.sigma_c;
    const float sn2 = 2.0f * bp.sigma_n * bp.sigma_n, sa2 = 2.0f * bp.sigma_a * bp.sigma_a;
    float wg = 0.0f, gr = 0.0f, gg = 0.0f, gb = 0.0f;
    float wc = 0.0f, cr = 0.0f, cg = 0.0f, cb = 0.0f;
    float wj = 0.0f, jr = 0.0f, jg = 0.0f, jb = 0.0f;
    for (int y = i - r; y < i + r + 1; y++) {
        const int c = min(max(y, 0), H - 1);
        for (int x = j - r; x < j + r + 1; x++) {
            const int col = min(max(x, 0), W - 1);
            const BlPix q = fetch(c, col);
            const float dsp = float((i - c) * (i - c) + (j - col) * (j - col));
            const float rd = ctr.r - q.r, gd = ctr.g - q.g, bd = ctr.b - q.b;
            float dsc = rd * rd + gd * gd + bd * bd;
            dsc *= 0.5f;
            const float ard = ctr.ar - q.ar, agd = ctr.ag - q.ag, abd = ctr.ab - q.ab;
            const float dsa = ard * ard + agd * agd + abd * abd;
            float dot = ctr.nx * q.nx + ctr.ny * q.ny + ctr.nz * q.nz;
            dot = fminf(fmaxf(dot, -1.0f), 1.0f);
            const float nv = acosf(dot);
            float nv2 = nv * nv;
            nv2 *= 1.5f;
            const float w_g = expf(-dsp / sp2);
            gr += w_g * q.r; gg += w_g * q.g; gb += w_g * q.b;
            const float w_c = expf(-dsp / sp2 - dsc / sc2);
            cr += w_c * q.r; cg += w_c * q.g; cb += w_c * q.b;
            const float w_j = expf(-dsp / sp2 - dsc / sc2 - nv2 / sn2 - dsa / sa2);
            jr += w_j * q.r; jg += w_j * q.g; jb += w_j * q.b;
            wg += w_g; wc += w_c; wj += w_j;
        }
    }
    const int64_t p = 3 * (int64_t(i) * W + j);
    if (bp.gaussian) { bp.gaussian[p] = gr / wg; bp.gaussian[p + 1] = gg / wg; bp.gaussian[p + 2] = gb / wg; }
    if (bp.bilateral) { bp.bilateral[p] = cr / wc; bp.bilateral[p + 1] = cg / wc; bp.bilateral[p + 2] = cb / wc; }
    if (bp.joint) { bp.joint[p] = jr / wj; bp.joint[p + 1] = jg / wj; bp.joint[p + 2] = jb / wj; }
}

// unit normals (normalised as stored) and the colour's luminance, per pixel
__global__ void k_atrous_prepare(const float* normal, int64_t n, float* unit) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float x = normal[3 * i], y = normal[3 * i + 1], z = normal[3 * i + 2];
    normalize_in_place(x, y, z);
    unit[3 * i] = x; unit[3 * i + 1] = y; unit[3 * i + 2] = z;
}

__device__ __forceinline__ float lum709(float r, float g, float b) { return 0.2126f * r + 0.7152f * g + 0.0722f * b; }

__global__ void k_atrous_luminance(const float* in, int64_t n, float* lum) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    lum[i] = lum709(in[3 * i], in[3 * i + 1], in[3 * i + 2]);
}

__global__ void __launch_bounds__(256) k_atrous_pass(AtrousPass ap) {
    const int x = int(blockIdx.x) * 16 + int(threadIdx.x) % 16, y = int(blockIdx.y) * 16 + int(threadIdx.x) / 16;
    const int W = ap.width, H = ap.height;
    if (x >= W || y >= H) return;
    const int64_t idx = int64_t(y) * W + x;
    const float c_r = ap.in[3 * idx], c_g = ap.in[3 * idx + 1], c_b = ap.in[3 * idx + 2];
    const float c_lum = ap.lum[idx];
    const float n_r = ap.unit[3 * idx], n_g = ap.unit[3 * idx + 1], n_b = ap.unit[3 * idx + 2];
    const float a_r = ap.albedo[3 * idx], a_g = ap.albedo[3 * idx + 1], a_b = ap.albedo[3 * idx + 2];
    const float sc2 = 2.0f * ap.sigma_c * ap.sigma_c, sn2 = 2.0f * ap.sigma_n * ap.sigma_n;
    const float sa2 = 2.0f * ap.sigma_a * ap.sigma_a;
    const float hk[5] = {1 / 16.0f, 1 / 4.0f, 3 / 8.0f, 1 / 4.0f, 1 / 16.0f};
    float sr = 0.0f, sg = 0.0f, sb = 0.0f, sw = 0.0f;
    for (int ky = -2; ky <= 2; ++ky) {
        const int ny = min(max(y + ky * ap.step, 0), H - 1);
        for (int kx = -2; kx <= 2; ++kx) {
            const int nx = min(max(x + kx * ap.step, 0), W - 1);
            const int64_t ni = int64_t(ny) * W + nx;
            const float p_r = ap.in[3 * ni], p_g = ap.in[3 * ni + 1], p_b = ap.in[3 * ni + 2];
            const float w_sp = hk[ky + 2] * hk[kx + 2];
            const float ld = c_lum - ap.lum[ni];
            const float w_c = expf(-ld * ld / sc2);
            float dn = n_r * ap.unit[3 * ni] + n_g * ap.unit[3 * ni + 1] + n_b * ap.unit[3 * ni + 2];
            dn = fminf(fmaxf(dn, -1.0f), 1.0f);
            const float ang = acosf(dn);
            const float w_n = expf(-ang * ang / sn2);
            const float dr = a_r - ap.albedo[3 * ni], dg = a_g - ap.albedo[3 * ni + 1], db = a_b - ap.albedo[3 * ni + 2];
            const float w_a = expf(-(dr * dr + dg * dg + db * db) / sa2);
            const float wt = w_sp * w_c * w_n * w_a;
            sr += wt * p_r; sg += wt * p_g; sb += wt * p_b; sw += wt;
        }
    }
    if (sw > RPD_EPS_SMALL) {
        ap.out[3 * idx] = sr / sw; ap.out[3 * idx + 1] = sg / sw; ap.out[3 * idx + 2] = sb / sw;
    } else {
        ap.out[3 * idx] = c_r; ap.out[3 * idx + 1] = c_g; ap.out[3 * idx + 2] = c_b;
    }
}

}  // namespace

int bilateral_radius(float sigma_p) {
    const int r = int(ceilf(3.0f * sigma_p));
    return r < 1 ? 1 : r;
}

void launch_joint_bilateral(const BilateralParams& bp, hipStream_t st) {
    const dim3 grid((bp.width + BL_TILE - 1) / BL_TILE, (bp.height + BL_TILE - 1) / BL_TILE);
    if (bp.radius <= BL_LDS_RADIUS) {
        const int span = BL_TILE + 2 * bp.radius;
        const size_t lds = size_t(span) * span * 9 * sizeof(float);
        hipLaunchKernelGGL(k_joint_bilateral<true>, grid, dim3(BL_TILE * BL_TILE), lds, st, bp);
    } else {
        hipLaunchKernelGGL(k_joint_bilateral<false>, grid, dim3(BL_TILE * BL_TILE), 0, st, bp);
    }
}

void launch_atrous_prepare(const float* normal, int64_t n, float* unit, hipStream_t st) {
    hipLaunchKernelGGL(k_atrous_prepare, dim3(unsigned((n + 255) / 256)), dim3(256), 0, st, normal, n, unit);
}

void launch_atrous_pass(const AtrousPass& ap, hipStream_t st) {
    const int64_t n = int64_t(ap.width) * ap.height;
    hipLaunchKernelGGL(k_atrous_luminance, dim3(unsigned((n + 255) / 256)), dim3(256), 0, st, ap.in, n, ap.lum);
    const dim3 grid((ap.width + 15) / 16, (ap.height + 15) / 16);
    hipLaunchKernelGGL(k_atrous_pass, grid, dim3(256), 0, st, ap);
}

}  // namespace rpd
