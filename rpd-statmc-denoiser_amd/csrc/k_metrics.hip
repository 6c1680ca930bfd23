// Image-quality metrics of the reference's evaluation tool on the device:
// hdr_metrics (shared/tools/eval_tools.cpp:215-285) and
// compute_extended_metrics (eval_tools.cpp:120-213), plus relMSE.
//
// Two passes over the pair of images, each followed by a fixed-order
// reduction of per-block partials (so a result does not depend on timing):
//   A: errors, energies, channel means, relMSE, log-luminance gradients and
//      the two maxima (reference peak, log-luminance peak);
//   B: the global per-channel SSIM moments about the means of A, and the
//      11x11 windowed SSIM of log luminance, whose constants need the
//      log-luminance peak of A. Each block stages the log luminance of its
//      16x16 tile and a 5-pixel halo in LDS and sums every window from there.
// All arithmetic is FP64, as in the reference. HBM-bound: 24 B per pixel and
// pass for interleaved float RGB.
#include "kernels.h"

namespace rpd {

namespace {

constexpr int MA_BLOCK = 256;
constexpr int MA_MAX_BLOCKS = METRICS_A_MAX_BLOCKS;
constexpr int MB_TILE = 16;
constexpr int MB_HALO = 5;
constexpr int MB_SPAN = MB_TILE + 2 * MB_HALO;   // 26

__device__ __forceinline__ double ld(const MetricImage& im, int64_t i, int c) { return double(im.p[c * im.cs + i * im.ps]); }

__device__ __forceinline__ double log_lum(const MetricImage& im, int64_t i) {
    const double l = 0.2126 * ld(im, i, 0) + 0.7152 * ld(im, i, 1) + 0.0722 * ld(im, i, 2);
    return log1p(fmax(0.0, l));
}

// Block-wide reduction of `v` (sum, or max when is_max) into s[0]; every
// thread gets the result. `s` holds MA_BLOCK doubles.
__device__ double block_reduce(double v, double* s, bool is_max) {
    s[threadIdx.x] = v;
    __syncthreads();
    for (int o = MA_BLOCK / 2; o > 0; o >>= 1) {
        if (threadIdx.x < o) s[threadIdx.x] = is_max ? fmax(s[threadIdx.x], s[threadIdx.x + o])
                                                     : s[threadIdx.x] + s[threadIdx.x + o];
        __syncthreads();
    }
    const double r = s[0];
    __syncthreads();
    return r;
}

__global__ void __launch_bounds__(MA_BLOCK) k_metrics_a(MetricImage ref, MetricImage tst, int W, int H, double* part) {
    __shared__ double s[MA_BLOCK];
    double acc[MA_FIELDS];
    for (int k = 0; k < MA_FIELDS; ++k) acc[k] = 0.0;
    const int64_t N = int64_t(W) * H;
    for (int64_t i = int64_t(blockIdx.x) * MA_BLOCK + threadIdx.x; i < N; i += int64_t(gridDim.x) * MA_BLOCK) {
        for (int c = 0; c < 3; ++c) {
            const double a = ld(ref, i, c), b = ld(tst, i, c);
            const double d = a - b;
            acc[MA_SUM_SQ] += d * d;
            acc[MA_SUM_ABS] += fabs(d);
            acc[MA_MEAN_REF + c] += a;
            acc[MA_MEAN_TST + c] += b;
            acc[MA_REF_ENERGY] += a * a;
            acc[MA_RELMSE] += (b - a) * (b - a) / (a * a + 1e-2);
            acc[MA_PEAK] = fmax(acc[MA_PEAK], fabs(a));
        }
        const int x = int(i % W), y = int(i / W);
        const double lr = log_lum(ref, i), lt = log_lum(tst, i);
        acc[MA_LOG_PEAK] = fmax(acc[MA_LOG_PEAK], lr);
        if (x < W - 1 && y < H - 1) {  // eval_tools.cpp:193-211: forward differences inside the image
            const double rdx = log_lum(ref, i + 1) - lr, rdy = log_lum(ref, i + W) - lr;
            const double tdx = log_lum(tst, i + 1) - lt, tdy = log_lum(tst, i + W) - lt;
            acc[MA_GRAD_ERR] += (rdx - tdx) * (rdx - tdx) + (rdy - tdy) * (rdy - tdy);
            acc[MA_GRAD_REF] += rdx * rdx + rdy * rdy;
        }
    }
    for (int k = 0; k < MA_FIELDS; ++k) {
        const double r = block_reduce(acc[k], s, k >= MA_PEAK);
        if (threadIdx.x == 0) part[int64_t(blockIdx.x) * MA_FIELDS + k] = r;
    }
}

__global__ void __launch_bounds__(MA_BLOCK) k_metrics_b(MetricImage ref, MetricImage tst, int W, int H,
                                                        MetricsPassB pb, double* part) {
    __shared__ double s_r[MB_SPAN * MB_SPAN];
    __shared__ double s_t[MB_SPAN * MB_SPAN];
    __shared__ double s[MA_BLOCK];
    const int tx0 = int(blockIdx.x) * MB_TILE, ty0 = int(blockIdx.y) * MB_TILE;
    for (int k = threadIdx.x; k < MB_SPAN * MB_SPAN; k += MA_BLOCK) {
        const int gx = tx0 - MB_HALO + k % MB_SPAN, gy = ty0 - MB_HALO + k / MB_SPAN;
        double a = 0.0, b = 0.0;
        if (gx >= 0 && gx < W && gy >= 0 && gy < H) {
            const int64_t i = int64_t(gy) * W + gx;
            a = log_lum(ref, i);
            b = log_lum(tst, i);
        }
        s_r[k] = a;
        s_t[k] = b;
    }
    __syncthreads();
    double acc[MB_FIELDS];
    for (int k = 0; k < MB_FIELDS; ++k) acc[k] = 0.0;
    const int x = tx0 + int(threadIdx.x) % MB_TILE, y = ty0 + int(threadIdx.x) / MB_TILE;
    if (x < W && y < H) {
        const int64_t i = int64_t(y) * W + x;
        for (int c = 0; c < 3; ++c) {  // eval_tools.cpp:259-267: moments about the channel means
            const double a = ld(ref, i, c) - pb.mean_ref[c], b = ld(tst, i, c) - pb.mean_tst[c];
            acc[MB_VAR_REF + c] += a * a;
            acc[MB_VAR_TST + c] += b * b;
            acc[MB_COV + c] += a * b;
        }
        // eval_tools.cpp:176-190: the window [x-5, x+5] x [y-5, y+5] clipped to the image
        const int x0 = max(0, x - MB_HALO), x1 = min(W, x + MB_HALO + 1);
        const int y0 = max(0, y - MB_HALO), y1 = min(H, y + MB_HALO + 1);
        double sr = 0.0, st = 0.0, srr = 0.0, stt = 0.0, srt = 0.0;
        for (int yy = y0; yy < y1; ++yy) {
            const int row = (yy - ty0 + MB_HALO) * MB_SPAN - tx0 + MB_HALO;
            for (int xx = x0; xx < x1; ++xx) {
                const double a = s_r[row + xx], b = s_t[row + xx];
                sr += a; st += b; srr += a * a; stt += b * b; srt += a * b;
            }
        }
        const double n = double((x1 - x0) * (y1 - y0));
        const double mr = sr / n, mt = st / n;
        const double vr = fmax(0.0, srr / n - mr * mr), vt = fmax(0.0, stt / n - mt * mt);
        const double cv = srt / n - mr * mt;
        const double num = (2.0 * mr * mt + pb.c1) * (2.0 * cv + pb.c2);
        const double den = (mr * mr + mt * mt + pb.c1) * (vr + vt + pb.c2);
        acc[MB_SSIM] = den > 0.0 ? num / den : 1.0;
    }
    const int64_t b = int64_t(blockIdx.y) * gridDim.x + blockIdx.x;
    for (int k = 0; k < MB_FIELDS; ++k) {
        const double r = block_reduce(acc[k], s, false);
        if (threadIdx.x == 0) part[b * MB_FIELDS + k] = r;
    }
}

// One block: field k of out = sum (or max for k >= max_from) over the
// partials of all blocks, in a fixed order.
__global__ void __launch_bounds__(MA_BLOCK) k_metrics_reduce(const double* part, int64_t n_blocks, int fields,
                                                             int max_from, double* out) {
    __shared__ double s[MA_BLOCK];
    for (int k = 0; k < fields; ++k) {
        const bool is_max = k >= max_from;
        double v = 0.0;
        for (int64_t b = threadIdx.x; b < n_blocks; b += MA_BLOCK) {
            const double p = part[b * fields + k];
            v = is_max ? fmax(v, p) : v + p;
        }
        const double r = block_reduce(v, s, is_max);
        if (threadIdx.x == 0) out[k] = r;
    }
}

}  // namespace

int64_t metrics_a_blocks(int W, int H) {
    const int64_t n = (int64_t(W) * H + MA_BLOCK - 1) / MA_BLOCK;
    return n < MA_MAX_BLOCKS ? n : MA_MAX_BLOCKS;
}
int64_t metrics_b_blocks(int W, int H) {
    return int64_t((W + MB_TILE - 1) / MB_TILE) * ((H + MB_TILE - 1) / MB_TILE);
}

void launch_metrics_a(const MetricImage& ref, const MetricImage& tst, int W, int H, double* part, double* out,
                      hipStream_t st) {
    const int64_t nb = metrics_a_blocks(W, H);
    hipLaunchKernelGGL(k_metrics_a, dim3(unsigned(nb)), dim3(MA_BLOCK), 0, st, ref, tst, W, H, part);
    hipLaunchKernelGGL(k_metrics_reduce, dim3(1), dim3(MA_BLOCK), 0, st, part, nb, int(MA_FIELDS), int(MA_PEAK), out);
}

void launch_metrics_b(const MetricImage& ref, const MetricImage& tst, int W, int H, const MetricsPassB& pb,
                      double* part, double* out, hipStream_t st) {
    const dim3 grid((W + MB_TILE - 1) / MB_TILE, (H + MB_TILE - 1) / MB_TILE);
    hipLaunchKernelGGL(k_metrics_b, grid, dim3(MA_BLOCK), 0, st, ref, tst, W, H, pb, part);
    hipLaunchKernelGGL(k_metrics_reduce, dim3(1), dim3(MA_BLOCK), 0, st, part, metrics_b_blocks(W, H),
                       int(MB_FIELDS), int(MB_FIELDS), out);
}

}  // namespace rpd
