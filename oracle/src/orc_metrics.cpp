// ORACLE — test infrastructure only (see orc_common.h).
// Image-quality metrics of the reference's evaluation tool, restated on
// float RGB images (interleaved, row-major) in double arithmetic:
//   * hdr_metrics (shared/tools/eval_tools.cpp:215-285): MAE, RMSE, PSNR
//     against the reference's peak, global per-channel SSIM;
//   * compute_extended_metrics (eval_tools.cpp:120-213): NRMSE, LogSSIM11
//     (11x11 windows clipped at the border, on log1p of Rec.709 luminance,
//     through integral images as there), LogGradNRMSE (forward differences);
//   * relMSE = mean over pixels and channels of (x - r)^2 / (r^2 + 1e-2),
//     which the reference does not compute.
#include <algorithm>
#include <cmath>
#include <vector>

extern "C" {

// Output order of orc_image_metrics (and of the engine's rpd_image_metrics):
// mae, rmse, psnr, peak, ssim r, ssim g, ssim b, ssim mean, nrmse,
// log_ssim11, log_gradient_nrmse, relmse.
enum { OM_MAE, OM_RMSE, OM_PSNR, OM_PEAK, OM_SSIM_R, OM_SSIM_G, OM_SSIM_B, OM_SSIM_MEAN, OM_NRMSE, OM_LOG_SSIM11,
       OM_LOG_GRAD_NRMSE, OM_RELMSE, OM_COUNT };

static double box_sum(const std::vector<double>& s, int stride, int x0, int y0, int x1, int y1) {
    return s[y1 * stride + x1] - s[y0 * stride + x1] - s[y1 * stride + x0] + s[y0 * stride + x0];
}

void orc_image_metrics(const float* ref, const float* tst, int w, int h, double* out) {
    const int pixels = w * h;
    // eval_tools.cpp:235-254
    double sum_sq = 0.0, sum_abs = 0.0, peak = 0.0, rel = 0.0;
    double mean_ref[3] = {0, 0, 0}, mean_tst[3] = {0, 0, 0};
    for (int i = 0; i < pixels; ++i)
        for (int c = 0; c < 3; ++c) {
            const double a = ref[3 * i + c], b = tst[3 * i + c];
            sum_sq += (a - b) * (a - b);
            sum_abs += std::abs(a - b);
            peak = std::max(peak, std::abs(a));
            mean_ref[c] += a;
            mean_tst[c] += b;
            rel += (b - a) * (b - a) / (a * a + 1e-2);
        }
    for (int c = 0; c < 3; ++c) { mean_ref[c] /= pixels; mean_tst[c] /= pixels; }
    // eval_tools.cpp:256-275: one SSIM over the whole image per channel
    double ssim[3];
    if (peak <= 0.0) peak = 1.0;
    const double C1 = (0.01 * peak) * (0.01 * peak), C2 = (0.03 * peak) * (0.03 * peak);
    for (int c = 0; c < 3; ++c) {
        double vr = 0.0, vt = 0.0, cov = 0.0;
        for (int i = 0; i < pixels; ++i) {
            const double a = ref[3 * i + c], b = tst[3 * i + c];
            vr += (a - mean_ref[c]) * (a - mean_ref[c]);
            vt += (b - mean_tst[c]) * (b - mean_tst[c]);
            cov += (a - mean_ref[c]) * (b - mean_tst[c]);
        }
        vr /= pixels; vt /= pixels; cov /= pixels;
        const double num = (2.0 * mean_ref[c] * mean_tst[c] + C1) * (2.0 * cov + C2);
        const double den = (mean_ref[c] * mean_ref[c] + mean_tst[c] * mean_tst[c] + C1) * (vr + vt + C2);
        ssim[c] = den > 0.0 ? num / den : 0.0;
    }
    const double mse = sum_sq / (pixels * 3);
    const double rmse = std::sqrt(mse);
    out[OM_MAE] = sum_abs / (pixels * 3);
    out[OM_RMSE] = rmse;
    out[OM_PSNR] = rmse > 0.0 ? 20.0 * std::log10(peak / rmse) : INFINITY;
    out[OM_PEAK] = peak;
    for (int c = 0; c < 3; ++c) out[OM_SSIM_R + c] = ssim[c];
    out[OM_SSIM_MEAN] = (ssim[0] + ssim[1] + ssim[2]) / 3.0;
    out[OM_RELMSE] = rel / (pixels * 3.0);

    // eval_tools.cpp:120-145: log luminance and the NRMSE energies
    std::vector<double> rl(pixels), tl(pixels);
    double ref_energy = 0.0, err_energy = 0.0, log_peak = 0.0;
    const double wts[3] = {0.2126, 0.7152, 0.0722};
    for (int i = 0; i < pixels; ++i) {
        double ra = 0.0, tb = 0.0;
        for (int c = 0; c < 3; ++c) {
            const double a = ref[3 * i + c], b = tst[3 * i + c];
            ref_energy += a * a;
            err_energy += (a - b) * (a - b);
            ra += wts[c] * a;
            tb += wts[c] * b;
        }
        rl[i] = std::log1p(std::max(0.0, ra));
        tl[i] = std::log1p(std::max(0.0, tb));
        log_peak = std::max(log_peak, rl[i]);
    }
    out[OM_NRMSE] = std::sqrt(err_energy / std::max(1e-18, ref_energy));
    // eval_tools.cpp:147-190: integral images, 11x11 windows clipped at the border
    const int stride = w + 1;
    std::vector<double> s_r(stride * (h + 1), 0.0), s_t(s_r), s_rr(s_r), s_tt(s_r), s_rt(s_r);
    for (int y = 0; y < h; ++y) {
        double a1 = 0, b1 = 0, a2 = 0, b2 = 0, ab = 0;
        for (int x = 0; x < w; ++x) {
            const double a = rl[y * w + x], b = tl[y * w + x];
            a1 += a; b1 += b; a2 += a * a; b2 += b * b; ab += a * b;
            const int o = (y + 1) * stride + x + 1;
            s_r[o] = s_r[o - stride] + a1;
            s_t[o] = s_t[o - stride] + b1;
            s_rr[o] = s_rr[o - stride] + a2;
            s_tt[o] = s_tt[o - stride] + b2;
            s_rt[o] = s_rt[o - stride] + ab;
        }
    }
    log_peak = std::max(1e-6, log_peak);
    const double c1 = std::pow(0.01 * log_peak, 2.0), c2 = std::pow(0.03 * log_peak, 2.0);
    double ssim_sum = 0.0;
    for (int y = 0; y < h; ++y) {
        const int y0 = std::max(0, y - 5), y1 = std::min(h, y + 6);
        for (int x = 0; x < w; ++x) {
            const int x0 = std::max(0, x - 5), x1 = std::min(w, x + 6);
            const double n = double((x1 - x0) * (y1 - y0));
            const double mr = box_sum(s_r, stride, x0, y0, x1, y1) / n;
            const double mt = box_sum(s_t, stride, x0, y0, x1, y1) / n;
            const double vr = std::max(0.0, box_sum(s_rr, stride, x0, y0, x1, y1) / n - mr * mr);
            const double vt = std::max(0.0, box_sum(s_tt, stride, x0, y0, x1, y1) / n - mt * mt);
            const double cv = box_sum(s_rt, stride, x0, y0, x1, y1) / n - mr * mt;
            const double num = (2.0 * mr * mt + c1) * (2.0 * cv + c2);
            const double den = (mr * mr + mt * mt + c1) * (vr + vt + c2);
            ssim_sum += den > 0.0 ? num / den : 1.0;
        }
    }
    out[OM_LOG_SSIM11] = ssim_sum / double(pixels);
    // eval_tools.cpp:193-211
    double ge = 0.0, gr = 0.0;
    for (int y = 0; y < h - 1; ++y)
        for (int x = 0; x < w - 1; ++x) {
            const int i = y * w + x;
            const double rdx = rl[i + 1] - rl[i], rdy = rl[i + w] - rl[i];
            const double tdx = tl[i + 1] - tl[i], tdy = tl[i + w] - tl[i];
            ge += (rdx - tdx) * (rdx - tdx) + (rdy - tdy) * (rdy - tdy);
            gr += rdx * rdx + rdy * rdy;
        }
    out[OM_LOG_GRAD_NRMSE] = std::sqrt(ge / std::max(1e-18, gr));
}

}  // extern "C"
