// StatMC Welch-compatibility colour filter and sample-variance smoothing.
//
// Reference: path_tracer/main.cpp:957-1162 (statmc_denoise), 188-231
// (lookup_t_critical), 292-327 (per-pixel reliable components),
// core/stats.h:48-91 (shared reliability, relaxation, variance update),
// main.cpp:1166-1268 (variance_denoise).
#include "kernels.h"

#include <algorithm>

namespace rpd {

__device__ __forceinline__ float tile_lookup_f(const float* t, GridGeom g, int x, int y) {
    const float fx = float(x) / float(g.step), fy = float(y) / float(g.step);
    const int x0 = min(max(int(floorf(fx)), 0), g.nx - 1), y0 = min(max(int(floorf(fy)), 0), g.ny - 1);
    const int x1 = min(x0 + 1, g.nx - 1), y1 = min(y0 + 1, g.ny - 1);
    const float tx = clampf(fx - float(x0), 0.0f, 1.0f), ty = clampf(fy - float(y0), 0.0f, 1.0f);
    const float top = lerpf_(t[y0 * g.nx + x0], t[y0 * g.nx + x1], tx);
    const float bot = lerpf_(t[y1 * g.nx + x0], t[y1 * g.nx + x1], tx);
    return lerpf_(top, bot, ty);
}

// Two-sided t critical value (main.cpp:188-231). Tables for alpha 0.05 and
// 0.01 (df 1..30, interpolated), otherwise a normal quantile step function.
__constant__ float c_t05[30] = {12.706f, 4.303f, 3.182f, 2.776f, 2.571f, 2.447f, 2.365f, 2.306f, 2.262f, 2.228f,
                                2.201f, 2.179f, 2.160f, 2.145f, 2.131f, 2.120f, 2.110f, 2.101f, 2.093f, 2.086f,
                                2.080f, 2.074f, 2.069f, 2.064f, 2.060f, 2.056f, 2.052f, 2.048f, 2.045f, 2.042f};
__constant__ float c_t01[30] = {63.657f, 9.925f, 5.841f, 4.604f, 4.032f, 3.707f, 3.499f, 3.355f, 3.250f, 3.169f,
                                3.106f, 3.055f, 3.012f, 2.977f, 2.947f, 2.921f, 2.898f, 2.878f, 2.861f, 2.845f,
                                2.831f, 2.819f, 2.807f, 2.797f, 2.787f, 2.779f, 2.771f, 2.763f, 2.756f, 2.750f};

// table selector: 0 = alpha 0.05, 1 = alpha 0.01, 2 = normal quantile
__device__ __forceinline__ float t_critical(float df, int table, float z) {
    df = fmaxf(df, 1.0f);
    if (table < 2 && df < 30.5f) {
        const float* tab = table == 0 ? c_t05 : c_t01;
        const int lo = min(max(int(floorf(df)), 1), 30), hi = min(max(int(ceilf(df)), 1), 30);
        if (lo == hi) return tab[lo - 1];
        return lerpf_(tab[lo - 1], tab[hi - 1], (df - float(lo)) / float(hi - lo));
    }
    if (df < 30.0f) return z * sqrtf(df / fmaxf(1.0f, df - 1.0f));
    return z;
}

// Per-pixel filter inputs, packed as 7 float4 planes:
//   P0 = (unit normal xyz, mean depth; -1 marks a pixel without primary hits)
//   P1 = (mean albedo rgb, n)
//   P2 = (sqrt-domain colour mean rgb, clamped luminance variance)
//   P3 = (stabilised sqrt variance of the mean rgb, variance of the mean)
//   P4 = (colour mean rgb, 0)
//   P5 = (reliable components 0..3), P6 = (reliable components 4..5, 0, 0)
// followed by 2 double2 planes D0 = (s0, s1), D1 = (s2, 0): the pixel's term
// s_ch = v_ch^2 / max(n - 1, 1) of each channel's Welch-Satterthwaite
// denominator (v = P3 channel, n = P1.w), formed once per pixel instead of once
// per (pixel, neighbour) pair; the same FP64 operations give the same value.
constexpr int FILTER_PLANES = 7;
__device__ __forceinline__ double2* filter_ws_planes(const ImageDesc& d) {
    return reinterpret_cast<double2*>(d.filter_scratch + (4 + 4 * FILTER_PLANES) * size_t(d.N));
}

// sqrt-domain variance of the mean, floored (main.cpp:977-980)
__global__ void k_sqrt_var_mean(Batch b) {
    const int64_t item = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (item >= b.total[BK_PIXEL]) return;
    int64_t li;
    const ImageDesc& d = b.img[batch_find(b, BK_PIXEL, item, li)];
    const int i = int(li), n = d.N;
    const PixelBuffers& pb = d.pb;
    float* vm = d.filter_scratch;
    const int cnt = pb.n[i];
    for (int c = 0; c < 3; ++c) {
        const float sv = cnt <= 1 ? 0.0f : pb.sqrt_m2[c * n + i] / float(cnt - 1);
        vm[c * n + i] = fmaxf(sv / float(max(1, cnt)), RPD_EPS_SMALL);
    }
}

__global__ void k_filter_pack(Batch b) {
    const int64_t item = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (item >= b.total[BK_PIXEL]) return;
    int64_t li;
    const ImageDesc& d = b.img[batch_find(b, BK_PIXEL, item, li)];
    const int W = d.W, H = d.H, n = d.N, i = int(li);
    const FilterConfig fc = d.fc;
    const PixelBuffers& pb = d.pb;
    const Features& f = d.feat;
    const GridGeom g = d.g;
    const float* tiles = d.tiles;
    const float* grad = d.grad;
    const float* vm = d.filter_scratch;
    float4* P = reinterpret_cast<float4*>(d.filter_scratch + 4 * size_t(n));
    const int x = i % W, y = i / W;
    // main.cpp:984-990: max of the value and its 3x3 weighted local average
    double a0 = 0, a1 = 0, a2 = 0, ws = 0;
    for (int dy = -1; dy <= 1; ++dy) {
        const int yy = min(max(y + dy, 0), H - 1);
        for (int dx = -1; dx <= 1; ++dx) {
            const int xx = min(max(x + dx, 0), W - 1);
            const double wt = 1.0 / (1.0 + double(dx * dx + dy * dy));
            const int q = yy * W + xx;
            a0 += double(vm[q]) * wt; a1 += double(vm[n + q]) * wt; a2 += double(vm[2 * n + q]) * wt;
            ws += wt;
        }
    }
    const v3 avg = mk3(float(a0 / ws), float(a1 / ws), float(a2 / ws));
    const v3 svs = mk3(fmaxf(vm[i], avg.x), fmaxf(vm[n + i], avg.y), fmaxf(vm[2 * n + i], avg.z));
    const int cnt = pb.n[i];
    const float var2_max = fc.sigma_max * fc.sigma_max;
    float var = cnt > 1 ? pb.lum_m2[i] / float(cnt - 1) : 0.0f;
    if (!isfinite(var)) var = RPD_EPS_SMALL;
    var = clampf(var, RPD_EPS_SMALL, var2_max);
    const int ni = max(1, cnt);
    const float vmean = fmaxf(RPD_EPS_SMALL, var / float(ni));
    const bool hit = pb.hit_count[i] > 0;
    P[0 * n + i] = make_float4(f.normal[i], f.normal[n + i], f.normal[2 * n + i], hit ? f.depth[i] : -1.0f);
    P[1 * n + i] = make_float4(f.albedo[i], f.albedo[n + i], f.albedo[2 * n + i], float(ni));
    P[2 * n + i] = make_float4(pb.sqrt_mean[i], pb.sqrt_mean[n + i], pb.sqrt_mean[2 * n + i], var);
    P[3 * n + i] = make_float4(svs.x, svs.y, svs.z, vmean);
    {
        double ws3[3];
        for (int ch = 0; ch < 3; ++ch) {
            const double v = double(ch == 0 ? svs.x : ch == 1 ? svs.y : svs.z);
            ws3[ch] = (v * v) / fmax(double(ni - 1), 1.0);
        }
        double2* D = filter_ws_planes(d);
        D[i] = make_double2(ws3[0], ws3[1]);
        D[n + i] = make_double2(ws3[2], 0.0);
    }
    P[4 * n + i] = make_float4(pb.color_mean[i], pb.color_mean[n + i], pb.color_mean[2 * n + i], 0.0f);
    // main.cpp:313-327, 305-310: six reliable components attenuated by the gradient
    const int nn = g.nx * g.ny;
    const int src[6] = {SF_BRDF, SF_LENS, SF_LIGHT_UV, SF_LIGHT_SELECT, SF_ENV, SF_RR};
    const float gr = clampf(grad[i], 0.0f, 1.0f);
    float c[6];
    for (int k = 0; k < 6; ++k) {
        const float v = clampf(tile_lookup_f(tiles + src[k] * nn, g, x, y), 0.0f, 1.0f);
        c[k] = clampf(v * (1.0f - gr), 0.0f, 1.0f);
    }
    P[5 * n + i] = make_float4(c[0], c[1], c[2], c[3]);
    P[6 * n + i] = make_float4(c[4], c[5], 0.0f, 0.0f);
}

// The window filter, one thread per pixel (main.cpp:998-1127). FP64 per
// neighbour exactly as the reference: Welch t^2 per channel against the
// t critical value at Welch-Satterthwaite df, the maximum over channels,
// RPD relaxation, hard rejection above 25, weight exp(-q/2)/(1+d^2).
// 1 / (1 + d^2) for integer squared offsets d^2 <= STATMC_SP_MAX (window radius up to 24),
// formed once per block; larger windows compute it in the loop
constexpr int STATMC_SP_MAX = 2 * 24 * 24;

// The gate inputs of the window (P0: unit normal and depth, P1: albedo and n)
// of a tile and its halo, staged in LDS when they fit; g0 null reads them from
// the packed planes in global memory.
struct GateTile {
    const float4* g0;
    const float4* g1;
    int span_w;      // row pitch of the staged rectangle
    int ox, oy;      // image coordinates of its first element
};

// Two operations of the reference's per-channel loop are replaced by ones that
// give the same bits:
//  * the t critical value depends on df only through float(df), and is the
//    constant z once float(df) >= 30.5 (main.cpp:216-230). df = vs^2 / den is
//    divided in FP64 only when an FP32 comparison vs^2 >= 31 den (a few float
//    roundings, < 1e-6 relative, against a 1.6% margin) cannot show df >= 31;
//  * when every channel's critical value is the same tc, the channel maximum of
//    RN(t2_c / RN(tc tc)) is RN(max_c t2_c / RN(tc tc)): correctly rounded
//    division by one divisor is monotone. One FP64 division instead of three.
__device__ __forceinline__ void statmc_pixel(const ImageDesc& d, const int x, const int y, const double* sp_tab,
                                             const GateTile& gt) {
    const int W = d.W, H = d.H, n = d.N;
    const FilterConfig fc = d.fc;
    const float4* __restrict__ P = reinterpret_cast<const float4*>(d.filter_scratch + 4 * size_t(n));
    const PixelBuffers& pb = d.pb;
    float *den = d.den, *unc = d.unc, *alp = d.alp, *vtot = d.vtot, *veff = d.veff, *vrat = d.vrat;
    const int i = y * W + x;
    const float4 p0 = P[i];
    if (p0.w < 0.0f) {  // no primary hit: keep the mean
        const int cnt = pb.n[i];
        den[i] = pb.color_mean[i]; den[n + i] = pb.color_mean[n + i]; den[2 * n + i] = pb.color_mean[2 * n + i];
        const float var = cnt > 1 ? pb.lum_m2[i] / float(cnt - 1) : 0.0f;
        unc[i] = var / float(max(1, cnt));
        alp[i] = 1.0f;
        return;
    }
    int table = 2;
    if (fabsf(fmaxf(fminf(fc.compat_alpha, 0.999999f), RPD_EPS_SMALL) - 0.05f) < RPD_EPS) table = 0;
    else if (fabsf(fmaxf(fminf(fc.compat_alpha, 0.999999f), RPD_EPS_SMALL) - 0.01f) < RPD_EPS) table = 1;
    float z;
    {
        const float a = fmaxf(fminf(fc.compat_alpha, 0.999999f), RPD_EPS_SMALL);
        if (a <= 0.001f) z = 3.29f;
        else if (a <= 0.005f) z = 2.81f;
        else if (a <= 0.01f) z = 2.58f;
        else if (a <= 0.02f) z = 2.33f;
        else if (a <= 0.05f) z = 1.96f;
        else if (a <= 0.10f) z = 1.64f;
        else z = 1.28f;
    }
    const double zz = double(z) * double(z);
    const v3 ni = mk3(p0.x, p0.y, p0.z);
    const float di = p0.w;
    const float4 p1 = P[n + i], p2 = P[2 * n + i], p3 = P[3 * n + i], p4 = P[4 * n + i], p5 = P[5 * n + i],
                 p6 = P[6 * n + i];
    const v3 ai = mk3(p1.x, p1.y, p1.z);
    const int n_i = int(p1.w);
    const float var_i = p2.w;
    const float vm_i = p3.w;
    const float ci[6] = {p5.x, p5.y, p5.z, p5.w, p6.x, p6.y};
    vtot[i] = var_i;
    veff[i] = vm_i;
    vrat[i] = (var_i > RPD_EPS_SMALL) ? (vm_i * float(n_i)) / var_i : 0.0f;
    // each pixel's term of each channel's Welch-Satterthwaite denominator
    const double2* __restrict__ D = filter_ws_planes(d);
    double df_i[3];
    {
        const double2 s01 = D[i], s2 = D[n + i];
        df_i[0] = s01.x; df_i[1] = s01.y; df_i[2] = s2.x;
    }

    double ws = 0, wss = 0, wvs = 0, c0 = 0, c1 = 0, c2 = 0, accepted = 0;
    const int R = fc.radius;
    for (int dy = -R; dy <= R; ++dy) {
        const int yy = y + dy;
        if (yy < 0 || yy >= H) continue;
        for (int dx = -R; dx <= R; ++dx) {
            const int xx = x + dx;
            if (xx < 0 || xx >= W) continue;
            const int q = yy * W + xx;
            float4 q0, q1;
            if (gt.g0) {
                const int k = (yy - gt.oy) * gt.span_w + (xx - gt.ox);
                q0 = gt.g0[k];
                if (q0.w < 0.0f) continue;
                if (dot3(ni, mk3(q0.x, q0.y, q0.z)) < fc.normal_threshold) continue;
                if (relative_depth_delta(di, q0.w) > fc.depth_threshold) continue;
                q1 = gt.g1[k];
            } else {
                q0 = P[q];
                if (q0.w < 0.0f) continue;
                if (dot3(ni, mk3(q0.x, q0.y, q0.z)) < fc.normal_threshold) continue;
                if (relative_depth_delta(di, q0.w) > fc.depth_threshold) continue;
                q1 = P[n + q];
            }
            if (norm3(ai - mk3(q1.x, q1.y, q1.z)) > 0.12f) continue;
            const float4 q2 = P[2 * n + q], q3 = P[3 * n + q];
            const int n_j = int(q1.w);
            const double2 sj01 = D[q], sj2 = D[n + q];
            const double df_j[3] = {sj01.x, sj01.y, sj2.x};
            const float smi[3] = {p2.x, p2.y, p2.z}, smj[3] = {q2.x, q2.y, q2.z};
            const float vsi[3] = {p3.x, p3.y, p3.z}, vsj[3] = {q3.x, q3.y, q3.z};
            double t2[3], tcs[3];
            bool all_z = true;
#pragma unroll
            for (int ch = 0; ch < 3; ++ch) {
                const double vi = double(vsi[ch]), vj = double(vsj[ch]);
                const double vs = vi + vj;
                const double md = double(smi[ch] - smj[ch]);
                t2[ch] = md * md / fmax(double(RPD_EPS_SMALL), vs);
                const double den_df = df_i[ch] + df_j[ch];
                tcs[ch] = double(z);
                if (den_df > double(RPD_EPS_SMALL)) {
                    const float vsf = float(vs);
                    if (!(vsf * vsf >= 31.0f * float(den_df))) {
                        tcs[ch] = double(t_critical(float((vs * vs) / den_df), table, z));
                        all_z = false;
                    }
                } else {
                    tcs[ch] = double(t_critical(float(double(n_i + n_j - 2)), table, z));
                    all_z = false;
                }
            }
            double mx;
            if (all_z) {
                mx = fmax(fmax(t2[0], t2[1]), t2[2]) / zz;
            } else {
                mx = 0.0;
#pragma unroll
                for (int ch = 0; ch < 3; ++ch) mx = fmax(mx, t2[ch] / (tcs[ch] * tcs[ch]));
            }
            const float4 q5 = P[5 * n + q], q6 = P[6 * n + q];
            const float cj[6] = {q5.x, q5.y, q5.z, q5.w, q6.x, q6.y};
            float shared = 0.0f;
#pragma unroll
            for (int k = 0; k < 6; ++k) shared = fmaxf(shared, fminf(ci[k], cj[k]));
            shared = clampf(shared, 0.0f, 1.0f);
            const float adj = clampf(fc.rpd_scale * shared, 0.0f, 1.0f);
            const double qv = mx / (1.0 + double(adj));
            if (qv > 25.0) continue;
            const int d2 = dx * dx + dy * dy;
            const double sp = d2 <= STATMC_SP_MAX ? sp_tab[d2] : 1.0 / (1.0 + double(d2));
            const double wj = sp * exp(-0.5 * qv);
            const float4 q4 = P[4 * n + q];
            c0 += double(q4.x) * wj; c1 += double(q4.y) * wj; c2 += double(q4.z) * wj;
            ws += wj;
            wss += wj * wj;
            wvs += wj * wj * double(q3.w);
            if ((dx != 0 || dy != 0) && wj > 0.05 * sp) accepted += 1.0;
        }
    }
    v3 mn = mk3(p4.x, p4.y, p4.z);
    double vn = double(vm_i);
    if (ws > double(RPD_EPS_SMALL)) {
        mn = mk3(float(c0 / ws), float(c1 / ws), float(c2 / ws));
        vn = wvs / (ws * ws);
    }
    const double neff = ws * ws / fmax(double(RPD_EPS_SMALL), wss);
    const double blend = (accepted >= 2.0 && ws >= 1.1) ? fmin(fmax(1.0 - 1.0 / sqrt(neff), 0.0), 0.8) : 0.0;
    const float a = float(1.0 - blend);
    const float bf = float(blend);
    den[i] = p4.x + (mn.x - p4.x) * bf;
    den[n + i] = p4.y + (mn.y - p4.y) * bf;
    den[2 * n + i] = p4.z + (mn.z - p4.z) * bf;
    unc[i] = float(double(a) * double(a) * double(vm_i) + blend * blend * vn +
                   2.0 * double(a) * blend * (1.0 / fmax(ws, double(RPD_EPS_SMALL))) * double(vm_i));
    alp[i] = a;
    if (d.dbg) {
        d.dbg[i] = float(accepted);
        d.dbg[n + i] = float(ws);
        d.dbg[2 * n + i] = float(neff);
    }
}

// Tile shapes of the window kernel: TW x TH pixels, one thread each. Every
// shape computes every pixel with the same operations, so all give the same
// bits; which runs fastest depends on the window radius and the image sizes,
// and the engine's tuner picks one by timing them.
const int STATMC_SHAPES[STATMC_SHAPE_COUNT][2] = {{16, 16}, {32, 8}, {8, 32}, {64, 4}, {16, 8}, {8, 16}, {8, 8}};

// LDS for the staged gate inputs of one tile and halo: two float4 per pixel.
__host__ __device__ inline size_t statmc_gate_bytes(int tw, int th, int radius) {
    return size_t(tw + 2 * radius) * size_t(th + 2 * radius) * 2 * sizeof(float4);
}
constexpr size_t STATMC_GATE_LDS_MAX = 48 * 1024;   // larger halos read the gate inputs from global memory

// Persistent StatMC: blocks take TW x TH tiles of any image of the batch from
// the XCD-partitioned work ranges; one thread per pixel of the tile. With
// `stage`, the block first copies the gate inputs of the tile and its halo
// (every radius of the batch fits in the dynamic LDS it was given) to LDS.
template <int TW, int TH>
__global__ void __launch_bounds__(TW * TH) k_statmc(Batch b, int stage) {
    __shared__ int64_t s_item;
    __shared__ double s_sp[STATMC_SP_MAX + 1];
    extern __shared__ float4 s_gate[];
    for (int k = threadIdx.x; k <= STATMC_SP_MAX; k += blockDim.x) s_sp[k] = 1.0 / (1.0 + double(k));
    int cursor = 0;
    const int tx = threadIdx.x % TW, ty = threadIdx.x / TW;
    while (true) {
        if (threadIdx.x == 0) s_item = batch_take(b, BK_TILE, cursor);
        __syncthreads();
        const int64_t item = s_item;
        __syncthreads();
        if (item < 0) break;
        int64_t lt;
        const ImageDesc& d = b.img[batch_find(b, BK_TILE, item, lt)];
        const int tiles_x = (d.W + TW - 1) / TW;
        const int x0 = int(lt % tiles_x) * TW, y0 = int(lt / tiles_x) * TH;
        const int x = x0 + tx, y = y0 + ty;
        GateTile gt;
        gt.g0 = nullptr; gt.g1 = nullptr; gt.span_w = 0; gt.ox = 0; gt.oy = 0;
        if (stage) {
            const int r = d.fc.radius;
            const int sw = TW + 2 * r, sh = TH + 2 * r, cnt = sw * sh;
            const float4* P = reinterpret_cast<const float4*>(d.filter_scratch + 4 * size_t(d.N));
            for (int k = threadIdx.x; k < cnt; k += TW * TH) {
                const int gx = x0 - r + k % sw, gy = y0 - r + k / sw;
                if (gx >= 0 && gx < d.W && gy >= 0 && gy < d.H) {
                    const int q = gy * d.W + gx;
                    s_gate[k] = P[q];
                    s_gate[cnt + k] = P[d.N + q];
                }
            }
            __syncthreads();
            gt.g0 = s_gate; gt.g1 = s_gate + cnt; gt.span_w = sw; gt.ox = x0 - r; gt.oy = y0 - r;
        }
        if (x < d.W && y < d.H) statmc_pixel(d, x, y, s_sp, gt);
        if (stage) __syncthreads();   // the next tile's copy overwrites the staged inputs
    }
}

void launch_statmc_prepare(const Batch& b, hipStream_t st) {
    const int64_t np = b.total[BK_PIXEL];
    hipLaunchKernelGGL(k_sqrt_var_mean, dim3((np + 255) / 256), dim3(256), 0, st, b);
    hipLaunchKernelGGL(k_filter_pack, dim3((np + 255) / 256), dim3(256), 0, st, b);
}

template <int TW, int TH>
static void launch_statmc_t(const Batch& b, hipStream_t st) {
    // resident blocks over the device (a function-local static is initialised once
    // even when several job threads launch at the same time), counted without the
    // staged inputs' LDS; a launch that stages them keeps to what fits beside it
    static const int cap = device_resident_blocks(reinterpret_cast<const void*>(&k_statmc<TW, TH>), TW * TH);
    const size_t lds = statmc_gate_bytes(TW, TH, b.max_radius);
    const int stage = b.max_radius > 0 && lds <= STATMC_GATE_LDS_MAX ? 1 : 0;
    int per_cu = 0, dev = 0, cus = 1;
    if (stage && hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, reinterpret_cast<const void*>(&k_statmc<TW, TH>),
                                                              TW * TH, lds) == hipSuccess && per_cu > 0) {
        hipGetDevice(&dev);
        if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0) cus = 1;
    } else {
        per_cu = 0;
    }
    const int64_t limit = stage && per_cu > 0 ? int64_t(per_cu) * cus : cap;
    const int64_t blocks = std::min<int64_t>(b.total[BK_TILE], limit);
    if (blocks <= 0) return;
    hipLaunchKernelGGL((k_statmc<TW, TH>), dim3(unsigned(blocks)), dim3(TW * TH), stage ? lds : 0, st, b, stage);
}

int statmc_shape_of(int tile_w, int tile_h) {
    for (int k = 0; k < STATMC_SHAPE_COUNT; ++k)
        if (STATMC_SHAPES[k][0] == tile_w && STATMC_SHAPES[k][1] == tile_h) return k;
    return -1;
}

void launch_statmc_window(const Batch& b, hipStream_t st) {
    switch (statmc_shape_of(b.tile_w, b.tile_h)) {
        case 1: launch_statmc_t<32, 8>(b, st); break;
        case 2: launch_statmc_t<8, 32>(b, st); break;
        case 3: launch_statmc_t<64, 4>(b, st); break;
        case 4: launch_statmc_t<16, 8>(b, st); break;
        case 5: launch_statmc_t<8, 16>(b, st); break;
        case 6: launch_statmc_t<8, 8>(b, st); break;
        default: launch_statmc_t<16, 16>(b, st); break;
    }
}

// ---------------------------------------------------------------------------
// Sample-variance smoothing (main.cpp:1166-1268)
__global__ void k_var_init(int n, PixelBuffers pb, float maxv, float* cur) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int cnt = pb.n[i];
    float v = cnt > 1 ? pb.lum_m2[i] / float(cnt - 1) : 0.0f;
    if (!isfinite(v)) v = RPD_EPS_SMALL;
    cur[i] = clampf(v, RPD_EPS_SMALL, maxv);
}
__global__ void __launch_bounds__(SMOOTH_TW * SMOOTH_TH) k_var_smooth(int W, int H, VarianceConfig vc,
                                                                      PixelBuffers pb, Features f, const float* src,
                                                                      float* dst, int staged) {
    extern __shared__ float4 s_tile[];
    const int R = vc.radius;
    const SmoothSrc S = smooth_stage(W, H, R, pb, f, src, staged ? s_tile : nullptr);
    const int x = int(blockIdx.x) * SMOOTH_TW + int(threadIdx.x) % SMOOTH_TW;
    const int y = int(blockIdx.y) * SMOOTH_TH + int(threadIdx.x) / SMOOTH_TW;
    if (x >= W || y >= H) return;
    const int i = y * W + x;
    float4 gi;
    float vsrc;
    int hi;
    S.get(x, y, gi, vsrc, hi);
    if (hi == 0) { dst[i] = vsrc; return; }
    const v3 ni = mk3(gi.x, gi.y, gi.z);
    const float di = gi.w;
    const float vi = fmaxf(RPD_EPS_SMALL, vsrc);
    float ws = 0.0f, wsum = 0.0f;
    for (int dy = -R; dy <= R; ++dy) {
        const int yy = y + dy;
        if (yy < 0 || yy >= H) continue;
        for (int dx = -R; dx <= R; ++dx) {
            const int xx = x + dx;
            if (xx < 0 || xx >= W || (dx == 0 && dy == 0)) continue;
            float4 gq;
            float vq;
            int hq;
            S.get(xx, yy, gq, vq, hq);
            if (hq == 0) continue;
            if (dot3(ni, mk3(gq.x, gq.y, gq.z)) < vc.normal_threshold) continue;
            if (relative_depth_delta(di, gq.w) > vc.depth_threshold) continue;
            const float vj = fmaxf(RPD_EPS_SMALL, vq);
            const float rd = fabsf(vi - vj) / fmaxf(RPD_EPS_SMALL, sqrtf(vi * vj));
            if (rd > vc.compat_sigma) continue;
            const float wj = 1.0f / (1.0f + float(dx * dx + dy * dy));
            wsum += wj * vj;
            ws += wj;
        }
    }
    float nm = vi;
    if (ws > RPD_EPS_SMALL) nm = wsum / ws;
    // core/stats.h:85-91
    const float blend = clampf(fmaxf(0.0f, nm - vi) / fmaxf(RPD_EPS_SMALL, vi + nm + vc.shrinkage_k), 0.0f, 1.0f);
    dst[i] = lerpf_(vi, nm, blend);
}
__global__ void k_var_final(int n, PixelBuffers pb, float maxv, const float* cur, float* out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int cnt = pb.n[i];
    float v = cnt > 1 ? pb.lum_m2[i] / float(cnt - 1) : 0.0f;
    if (!isfinite(v)) v = RPD_EPS_SMALL;
    v = clampf(v, RPD_EPS_SMALL, maxv);
    out[i] = fmaxf(cur[i], v);
}

void launch_variance(int W, int H, const VarianceConfig& vc, const PixelBuffers& pb, const Features& f, float* tmp,
                     float* out, hipStream_t st) {
    const int n = W * H;
    const float maxv = vc.sigma_max * vc.sigma_max;
    const dim3 grid((n + 255) / 256);
    // ping-pong between out and tmp; the result of the last iteration ends in `cur`
    float* cur = out;
    float* nxt = tmp;
    hipLaunchKernelGGL(k_var_init, grid, dim3(256), 0, st, n, pb, maxv, cur);
    const dim3 tiles((W + SMOOTH_TW - 1) / SMOOTH_TW, (H + SMOOTH_TH - 1) / SMOOTH_TH);
    const size_t lds = smooth_lds_bytes(vc.radius);
    for (int it = 0; it < vc.iterations; ++it) {
        hipLaunchKernelGGL(k_var_smooth, tiles, dim3(SMOOTH_TW * SMOOTH_TH), lds, st, W, H, vc, pb, f, cur, nxt,
                           lds > 0 ? 1 : 0);
        float* t = cur; cur = nxt; nxt = t;
    }
    hipLaunchKernelGGL(k_var_final, grid, dim3(256), 0, st, n, pb, maxv, cur, out);
}

}  // namespace rpd
