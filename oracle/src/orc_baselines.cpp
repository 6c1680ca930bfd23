// ORACLE — test infrastructure only (see orc_common.h).
// The reference's two comparison denoisers, restated in float arithmetic
// with the reference's loop and expression order, on interleaved RGB float
// images as its tools read them from .hdr files:
//   * denoising/joint_bilateral.cpp:15-128: Gaussian, colour bilateral and
//     joint (colour + normal angle + albedo) bilateral filters in one sweep,
//     window radius max(1, ceil(3 sigma_p)), coordinates clamped to the
//     image, normals decoded as 2n - 1 and normalised;
//   * denoising/atrous_wavelet.cpp:22-183: edge-avoiding a-trous passes with
//     the 5x5 B3-spline kernel, step 2^iteration, luminance / normal angle /
//     albedo edge stops, normals normalised as stored.
// exp and acos are the float overloads, which the reference's unqualified
// calls resolve to once stb's <math.h> has brought them into scope.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <utility>
#include <vector>

namespace {

constexpr float kEpsSmall = 1e-8f;  // shared/global.h:4

void normalize3(float& x, float& y, float& z) {
    const float l2 = x * x + y * y + z * z;
    if (l2 > kEpsSmall) {
        const float inv = 1.0f / std::sqrt(l2);
        x *= inv; y *= inv; z *= inv;
    }
}

float luminance(float r, float g, float b) { return 0.2126f * r + 0.7152f * g + 0.0722f * b; }

}  // namespace

extern "C" {

// joint_bilateral.cpp:15-128; any output pointer may be null.
void orc_joint_bilateral(const float* raw, const float* normal, const float* albedo, int w, int h, float sigma_c,
                         float sigma_n, float sigma_p, float sigma_a, float* gaussian, float* bilateral, float* joint) {
    const int radius = std::max(1, int(std::ceil(3.0f * sigma_p)));
    for (int i = 0; i < h; i++) {
        for (int j = 0; j < w; j++) {
            float wg = 0.0f, gr = 0.0f, gg = 0.0f, gb = 0.0f;
            float wc = 0.0f, cr = 0.0f, cg = 0.0f, cb = 0.0f;
            float wj = 0.0f, jr = 0.0f, jg = 0.0f, jb = 0.0f;
            const int p = 3 * (i * w + j);
            const float red = raw[p], green = raw[p + 1], blue = raw[p + 2];
            float nir = normal[p] * 2.0f - 1.0f, nig = normal[p + 1] * 2.0f - 1.0f, nib = normal[p + 2] * 2.0f - 1.0f;
            normalize3(nir, nig, nib);
            const float air = albedo[p], aig = albedo[p + 1], aib = albedo[p + 2];
            for (int y = i - radius; y < i + radius + 1; y++) {
                for (int x = j - radius; x < j + radius + 1; x++) {
                    const int c = int(std::min(std::max(0.0f, float(y)), float(h) - 1));
                    const int r = int(std::min(std::max(0.0f, float(x)), float(w) - 1));
                    const float alpha = 0.5f, beta = 2.0f - alpha;
                    const int q = 3 * (c * w + r);
                    const float dsp = float((i - c) * (i - c) + (j - r) * (j - r));
                    const float rd = red - raw[q], gd = green - raw[q + 1], bd = blue - raw[q + 2];
                    float dsc = rd * rd + gd * gd + bd * bd;
                    dsc *= alpha;
                    float njr = normal[q] * 2.0f - 1.0f, njg = normal[q + 1] * 2.0f - 1.0f, njb = normal[q + 2] * 2.0f - 1.0f;
                    normalize3(njr, njg, njb);
                    const float ard = air - albedo[q], agd = aig - albedo[q + 1], abd = aib - albedo[q + 2];
                    const float dsa = ard * ard + agd * agd + abd * abd;
                    float dot = nir * njr + nig * njg + nib * njb;
                    dot = std::clamp(dot, -1.0f, 1.0f);
                    const float nv = std::acos(dot);
                    float nv2 = nv * nv;
                    nv2 *= beta;
                    const float w_g = std::exp(-dsp / (2.0f * sigma_p * sigma_p));
                    gr += w_g * raw[q]; gg += w_g * raw[q + 1]; gb += w_g * raw[q + 2];
                    const float w_c = std::exp(-dsp / (2.0f * sigma_p * sigma_p) - dsc / (2.0f * sigma_c * sigma_c));
                    cr += w_c * raw[q]; cg += w_c * raw[q + 1]; cb += w_c * raw[q + 2];
                    const float w_j = std::exp(-dsp / (2.0f * sigma_p * sigma_p) - dsc / (2.0f * sigma_c * sigma_c) -
                                               nv2 / (2.0f * sigma_n * sigma_n) - dsa / (2.0f * sigma_a * sigma_a));
                    jr += w_j * raw[q]; jg += w_j * raw[q + 1]; jb += w_j * raw[q + 2];
                    wg += w_g; wc += w_c; wj += w_j;
                }
            }
            if (gaussian) { gaussian[p] = gr / wg; gaussian[p + 1] = gg / wg; gaussian[p + 2] = gb / wg; }
            if (bilateral) { bilateral[p] = cr / wc; bilateral[p + 1] = cg / wc; bilateral[p + 2] = cb / wc; }
            if (joint) { joint[p] = jr / wj; joint[p + 1] = jg / wj; joint[p + 2] = jb / wj; }
        }
    }
}

// atrous_wavelet.cpp:22-44 (kernel), 57-145 (one pass), 147-183 (iterations)
void orc_atrous(const float* raw, const float* normal, const float* albedo, int w, int h, float sigma_c,
                float sigma_n, float sigma_a, int iterations, float* out) {
    static const float hk[5] = {1 / 16.0f, 1 / 4.0f, 3 / 8.0f, 1 / 4.0f, 1 / 16.0f};
    const size_t n3 = size_t(w) * h * 3;
    std::vector<float> a(raw, raw + n3), b(n3);
    float* in = a.data();
    float* ot = b.data();
    for (int it = 0; it < iterations; it++) {
        const int step = 1 << it;
        for (int y = 0; y < h; y++) {
            for (int x = 0; x < w; x++) {
                const int idx = y * w + x;
                const float c_r = in[3 * idx], c_g = in[3 * idx + 1], c_b = in[3 * idx + 2];
                const float c_lum = luminance(c_r, c_g, c_b);
                float n_r = normal[3 * idx], n_g = normal[3 * idx + 1], n_b = normal[3 * idx + 2];
                normalize3(n_r, n_g, n_b);
                const float a_r = albedo[3 * idx], a_g = albedo[3 * idx + 1], a_b = albedo[3 * idx + 2];
                float sr = 0.0f, sg = 0.0f, sb = 0.0f, sw = 0.0f;
                for (int k = 0; k < 25; k++) {
                    const int kx = k % 5 - 2, ky = k / 5 - 2;     // offsets in row-major order, dy outer
                    const int nx = std::clamp(x + kx * step, 0, w - 1);
                    const int ny = std::clamp(y + ky * step, 0, h - 1);
                    const int ni = ny * w + nx;
                    const float p_r = in[3 * ni], p_g = in[3 * ni + 1], p_b = in[3 * ni + 2];
                    const float p_lum = luminance(p_r, p_g, p_b);
                    float pn_r = normal[3 * ni], pn_g = normal[3 * ni + 1], pn_b = normal[3 * ni + 2];
                    normalize3(pn_r, pn_g, pn_b);
                    const float w_sp = hk[ky + 2] * hk[kx + 2];
                    const float ld = c_lum - p_lum;
                    const float w_c = std::exp(-ld * ld / (2.0f * sigma_c * sigma_c));
                    float dn = n_r * pn_r + n_g * pn_g + n_b * pn_b;
                    dn = std::clamp(dn, -1.0f, 1.0f);
                    const float ang = std::acos(dn);
                    const float w_n = std::exp(-ang * ang / (2.0f * sigma_n * sigma_n));
                    const float dr = a_r - albedo[3 * ni], dg = a_g - albedo[3 * ni + 1], db = a_b - albedo[3 * ni + 2];
                    const float das = dr * dr + dg * dg + db * db;
                    const float w_a = std::exp(-das / (2.0f * sigma_a * sigma_a));
                    const float wt = w_sp * w_c * w_n * w_a;
                    sr += wt * p_r; sg += wt * p_g; sb += wt * p_b; sw += wt;
                }
                if (sw > kEpsSmall) {
                    ot[3 * idx] = sr / sw; ot[3 * idx + 1] = sg / sw; ot[3 * idx + 2] = sb / sw;
                } else {
                    ot[3 * idx] = c_r; ot[3 * idx + 1] = c_g; ot[3 * idx + 2] = c_b;
                }
            }
        }
        std::swap(in, ot);
    }
    std::memcpy(out, in, n3 * sizeof(float));
}

}  // extern "C"
