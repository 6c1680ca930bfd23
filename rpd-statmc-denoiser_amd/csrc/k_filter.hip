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

struct StatmcConst {
    int table;   // t-critical table: 0 (alpha 0.05), 1 (alpha 0.01), 2 (none)
    float z;
    double zz;   // z * z
};
__device__ __forceinline__ StatmcConst statmc_const(const FilterConfig& fc) {
    StatmcConst k;
    k.table = 2;
    if (fabsf(fmaxf(fminf(fc.compat_alpha, 0.999999f), RPD_EPS_SMALL) - 0.05f) < RPD_EPS) k.table = 0;
    else if (fabsf(fmaxf(fminf(fc.compat_alpha, 0.999999f), RPD_EPS_SMALL) - 0.01f) < RPD_EPS) k.table = 1;
    const float a = fmaxf(fminf(fc.compat_alpha, 0.999999f), RPD_EPS_SMALL);
    if (a <= 0.001f) k.z = 3.29f;
    else if (a <= 0.005f) k.z = 2.81f;
    else if (a <= 0.01f) k.z = 2.58f;
    else if (a <= 0.02f) k.z = 2.33f;
    else if (a <= 0.05f) k.z = 1.96f;
    else if (a <= 0.10f) k.z = 1.64f;
    else k.z = 1.28f;
    k.zz = double(k.z) * double(k.z);
    return k;
}

// The pair tests of main.cpp:1040-1062 on neighbour j's P0 (normal, depth) and
// P1 (albedo, n) against pixel i's normal ni, depth di and albedo ai. Each test
// gives the same result with i and j exchanged: the dot product's products
// commute and are summed in one order, |a - b| = |b - a| and (a - b)^2 =
// (b - a)^2 exactly, and max and min commute.
__device__ __forceinline__ bool statmc_gate0(const FilterConfig& fc, const v3 ni, const float di, const float4 q0) {
    if (q0.w < 0.0f) return false;
    if (dot3(ni, mk3(q0.x, q0.y, q0.z)) < fc.normal_threshold) return false;
    return !(relative_depth_delta(di, q0.w) > fc.depth_threshold);
}
__device__ __forceinline__ bool statmc_gate1(const v3 ai, const float4 q1) {
    return !(norm3(ai - mk3(q1.x, q1.y, q1.z)) > 0.12f);
}

// The relaxed Welch statistic q of a pair that passed the gates
// (main.cpp:1063-1097): pixel i's P2, P3, Welch-Satterthwaite terms and
// reliable components against neighbour j's. Every operation is symmetric in
// i and j (a difference only squared, commutative sums, max and min), so q is
// the same bits with the two exchanged.
//
// Two operations of the reference's per-channel loop are replaced by ones that
// give the same bits:
//  * the t critical value depends on df only through float(df), and is the
//    constant z once float(df) >= 30.5 (main.cpp:216-230). df = vs^2 / den is
//    divided in FP64 only when an FP32 comparison vs^2 >= 31 den (a few float
//    roundings, < 1e-6 relative, against a 1.6% margin) cannot show df >= 31;
//  * when every channel's critical value is the same tc, the channel maximum of
//    RN(t2_c / RN(tc tc)) is RN(max_c t2_c / RN(tc tc)): correctly rounded
//    division by one divisor is monotone. One FP64 division instead of three.
__device__ __forceinline__ double statmc_q(const FilterConfig& fc, const StatmcConst& K, const int n_i, const float4 p2,
                                           const float4 p3, const double (&df_i)[3], const float (&ci)[6],
                                           const int n_j, const float4 q2, const float4 q3, const double (&df_j)[3],
                                           const float4 q5, const float4 q6) {
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
        tcs[ch] = double(K.z);
        if (den_df > double(RPD_EPS_SMALL)) {
            const float vsf = float(vs);
            if (!(vsf * vsf >= 31.0f * float(den_df))) {
                tcs[ch] = double(t_critical(float((vs * vs) / den_df), K.table, K.z));
                all_z = false;
            }
        } else {
            tcs[ch] = double(t_critical(float(double(n_i + n_j - 2)), K.table, K.z));
            all_z = false;
        }
    }
    double mx;
    if (all_z) {
        mx = fmax(fmax(t2[0], t2[1]), t2[2]) / K.zz;
    } else {
        mx = 0.0;
#pragma unroll
        for (int ch = 0; ch < 3; ++ch) mx = fmax(mx, t2[ch] / (tcs[ch] * tcs[ch]));
    }
    const float cj[6] = {q5.x, q5.y, q5.z, q5.w, q6.x, q6.y};
    float shared = 0.0f;
#pragma unroll
    for (int k = 0; k < 6; ++k) shared = fmaxf(shared, fminf(ci[k], cj[k]));
    shared = clampf(shared, 0.0f, 1.0f);
    const float adj = clampf(fc.rpd_scale * shared, 0.0f, 1.0f);
    return mx / (1.0 + double(adj));
}

__device__ __forceinline__ double statmc_sp(const double* sp_tab, int d2) {
    return d2 <= STATMC_SP_MAX ? sp_tab[d2] : 1.0 / (1.0 + double(d2));
}

// A pixel's own filter inputs (the i side of every pair).
struct StatmcSelf {
    float4 p0, p1, p2, p3, p4, p5, p6;
    v3 ni, ai;
    float di;
    int n_i;
    float ci[6];
    double df_i[3];
};
__device__ __forceinline__ void statmc_self(const ImageDesc& d, const int i, StatmcSelf& S) {
    const int n = d.N;
    const float4* __restrict__ P = reinterpret_cast<const float4*>(d.filter_scratch + 4 * size_t(n));
    S.p0 = P[i];
    S.p1 = P[n + i]; S.p2 = P[2 * n + i]; S.p3 = P[3 * n + i]; S.p4 = P[4 * n + i]; S.p5 = P[5 * n + i];
    S.p6 = P[6 * n + i];
    S.ni = mk3(S.p0.x, S.p0.y, S.p0.z);
    S.di = S.p0.w;
    S.ai = mk3(S.p1.x, S.p1.y, S.p1.z);
    S.n_i = int(S.p1.w);
    S.ci[0] = S.p5.x; S.ci[1] = S.p5.y; S.ci[2] = S.p5.z; S.ci[3] = S.p5.w; S.ci[4] = S.p6.x; S.ci[5] = S.p6.y;
    const double2* __restrict__ D = filter_ws_planes(d);
    const double2 s01 = D[i], s2 = D[n + i];
    S.df_i[0] = s01.x; S.df_i[1] = s01.y; S.df_i[2] = s2.x;
}

// Weight of neighbour q of pixel i at offset (dx, dy) (main.cpp:1032-1097), or
// -1 when a gate or the Welch test rejects it; q0 and q1 are q's P0 and P1.
__device__ __forceinline__ double statmc_pair_weight(const ImageDesc& d, const StatmcConst& K, const StatmcSelf& S,
                                                     const double* sp_tab, const int q, const float4 q0,
                                                     const float4 q1, const int dx, const int dy) {
    const FilterConfig& fc = d.fc;
    if (!statmc_gate0(fc, S.ni, S.di, q0)) return -1.0;
    if (!statmc_gate1(S.ai, q1)) return -1.0;
    const int n = d.N;
    const float4* __restrict__ P = reinterpret_cast<const float4*>(d.filter_scratch + 4 * size_t(n));
    const double2* __restrict__ D = filter_ws_planes(d);
    const float4 q2 = P[2 * n + q], q3 = P[3 * n + q];
    const double2 sj01 = D[q], sj2 = D[n + q];
    const double df_j[3] = {sj01.x, sj01.y, sj2.x};
    const float4 q5 = P[5 * n + q], q6 = P[6 * n + q];
    const double qv = statmc_q(fc, K, S.n_i, S.p2, S.p3, S.df_i, S.ci, int(q1.w), q2, q3, df_j, q5, q6);
    if (qv > 25.0) return -1.0;
    return statmc_sp(sp_tab, dx * dx + dy * dy) * exp(-0.5 * qv);
}

// A pixel without a primary hit keeps its mean (main.cpp:1000-1010).
__device__ __forceinline__ void statmc_keep_mean(const ImageDesc& d, const int i) {
    const int n = d.N;
    const PixelBuffers& pb = d.pb;
    const int cnt = pb.n[i];
    d.den[i] = pb.color_mean[i]; d.den[n + i] = pb.color_mean[n + i]; d.den[2 * n + i] = pb.color_mean[2 * n + i];
    const float var = cnt > 1 ? pb.lum_m2[i] / float(cnt - 1) : 0.0f;
    d.unc[i] = var / float(max(1, cnt));
    d.alp[i] = 1.0f;
}

// The window's sums of one pixel, in the window's row-major order.
struct StatmcSums {
    double ws = 0, wss = 0, wvs = 0, c0 = 0, c1 = 0, c2 = 0, accepted = 0;
    __device__ __forceinline__ void add(const double wj, const double sp, const float4 q4, const float q3w,
                                        const bool centre) {
        c0 += double(q4.x) * wj; c1 += double(q4.y) * wj; c2 += double(q4.z) * wj;
        ws += wj;
        wss += wj * wj;
        wvs += wj * wj * double(q3w);
        if (!centre && wj > 0.05 * sp) accepted += 1.0;
    }
};

// The blend, uncertainty and diagnostics from the window's sums
// (main.cpp:1099-1126).
__device__ __forceinline__ void statmc_finish(const ImageDesc& d, const int i, const StatmcSelf& S, const StatmcSums& A) {
    const int n = d.N;
    const float var_i = S.p2.w, vm_i = S.p3.w;
    d.vtot[i] = var_i;
    d.veff[i] = vm_i;
    d.vrat[i] = (var_i > RPD_EPS_SMALL) ? (vm_i * float(S.n_i)) / var_i : 0.0f;
    const float4 p4 = S.p4;
    v3 mn = mk3(p4.x, p4.y, p4.z);
    double vn = double(vm_i);
    if (A.ws > double(RPD_EPS_SMALL)) {
        mn = mk3(float(A.c0 / A.ws), float(A.c1 / A.ws), float(A.c2 / A.ws));
        vn = A.wvs / (A.ws * A.ws);
    }
    const double neff = A.ws * A.ws / fmax(double(RPD_EPS_SMALL), A.wss);
    const double blend = (A.accepted >= 2.0 && A.ws >= 1.1) ? fmin(fmax(1.0 - 1.0 / sqrt(neff), 0.0), 0.8) : 0.0;
    const float a = float(1.0 - blend);
    const float bf = float(blend);
    d.den[i] = p4.x + (mn.x - p4.x) * bf;
    d.den[n + i] = p4.y + (mn.y - p4.y) * bf;
    d.den[2 * n + i] = p4.z + (mn.z - p4.z) * bf;
    d.unc[i] = float(double(a) * double(a) * double(vm_i) + blend * blend * vn +
                     2.0 * double(a) * blend * (1.0 / fmax(A.ws, double(RPD_EPS_SMALL))) * double(vm_i));
    d.alp[i] = a;
    if (d.dbg) {
        d.dbg[i] = float(A.accepted);
        d.dbg[n + i] = float(A.ws);
        d.dbg[2 * n + i] = float(neff);
    }
}

__device__ __forceinline__ void statmc_pixel(const ImageDesc& d, const int x, const int y, const double* sp_tab,
                                             const GateTile& gt) {
    const int W = d.W, H = d.H, n = d.N;
    const float4* __restrict__ P = reinterpret_cast<const float4*>(d.filter_scratch + 4 * size_t(n));
    const int i = y * W + x;
    if (P[i].w < 0.0f) {  // no primary hit: keep the mean
        statmc_keep_mean(d, i);
        return;
    }
    const StatmcConst K = statmc_const(d.fc);
    StatmcSelf S;
    statmc_self(d, i, S);
    StatmcSums A;
    const int R = d.fc.radius;
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
                q0 = gt.g0[k]; q1 = gt.g1[k];
            } else {
                q0 = P[q]; q1 = P[n + q];
            }
            const double wj = statmc_pair_weight(d, K, S, sp_tab, q, q0, q1, dx, dy);
            if (wj < 0.0) continue;
            A.add(wj, statmc_sp(sp_tab, dx * dx + dy * dy), P[4 * n + q], P[3 * n + q].w, dx == 0 && dy == 0);
        }
    }
    statmc_finish(d, i, S, A);
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
// the pair pass stages the tile and its forward halo only (no rows above)
__host__ __device__ inline size_t statmc_gate_bytes_forward(int tw, int th, int radius) {
    return size_t(tw + 2 * radius) * size_t(th + radius) * 2 * sizeof(float4);
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

// ---- the symmetric two-pass window -------------------------------------------
// A pair's weight is the same bits with its two pixels exchanged (the gates
// and q above; the spatial term depends on dx^2 + dy^2). So the pair pass
// computes each pair once, for the forward half of the window (dy > 0, or
// dy == 0 and dx > 0), into plane k of `pair_w` (plane k holds offset k's
// weight of every pixel: N doubles, -1 for a rejected or missing neighbour),
// and the sum pass walks each pixel's full window in the reference's order,
// taking a forward neighbour's weight from its own row, a backward one's from
// the neighbour's row, and computing the centre's. Half the pairs' FP64 work
// of the one-pass form, the same sums in the same order.
__device__ __forceinline__ int statmc_forward_index(int dx, int dy, int r) {
    return dy == 0 ? dx - 1 : r + (dy - 1) * (2 * r + 1) + (dx + r);
}
__device__ __forceinline__ void statmc_pairs_pixel(const ImageDesc& d, const int x, const int y, const double* sp_tab,
                                                   const GateTile& gt) {
    const int W = d.W, H = d.H, n = d.N, R = d.fc.radius;
    const float4* __restrict__ P = reinterpret_cast<const float4*>(d.filter_scratch + 4 * size_t(n));
    double* __restrict__ Wp = d.pair_w;
    const int i = y * W + x;
    if (P[i].w < 0.0f) {   // a neighbour without a primary hit is skipped either way
        for (int k = 0; k < statmc_forward_count(R); ++k) Wp[size_t(k) * n + i] = -1.0;
        return;
    }
    const StatmcConst K = statmc_const(d.fc);
    StatmcSelf S;
    statmc_self(d, i, S);
    for (int dy = 0; dy <= R; ++dy) {
        const int yy = y + dy;
        for (int dx = dy == 0 ? 1 : -R; dx <= R; ++dx) {
            const int xx = x + dx;
            double wj = -1.0;
            if (yy < H && xx >= 0 && xx < W) {
                const int q = yy * W + xx;
                float4 q0, q1;
                if (gt.g0) {
                    const int k = (yy - gt.oy) * gt.span_w + (xx - gt.ox);
                    q0 = gt.g0[k]; q1 = gt.g1[k];
                } else {
                    q0 = P[q]; q1 = P[n + q];
                }
                wj = statmc_pair_weight(d, K, S, sp_tab, q, q0, q1, dx, dy);
            }
            Wp[size_t(statmc_forward_index(dx, dy, R)) * n + i] = wj;
        }
    }
}

template <int TW, int TH>
__global__ void __launch_bounds__(TW * TH) k_statmc_pairs(Batch b, int stage) {
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
        if (stage) {   // the tile and its forward halo: rows y0 .. y0 + TH - 1 + r
            const int r = d.fc.radius;
            const int sw = TW + 2 * r, sh = TH + r, cnt = sw * sh;
            const float4* P = reinterpret_cast<const float4*>(d.filter_scratch + 4 * size_t(d.N));
            for (int k = threadIdx.x; k < cnt; k += TW * TH) {
                const int gx = x0 - r + k % sw, gy = y0 + k / sw;
                if (gx >= 0 && gx < d.W && gy < d.H) {
                    const int q = gy * d.W + gx;
                    s_gate[k] = P[q];
                    s_gate[cnt + k] = P[d.N + q];
                }
            }
            __syncthreads();
            gt.g0 = s_gate; gt.g1 = s_gate + cnt; gt.span_w = sw; gt.ox = x0 - r; gt.oy = y0;
        }
        if (x < d.W && y < d.H) statmc_pairs_pixel(d, x, y, s_sp, gt);
        if (stage) __syncthreads();   // the next tile's copy overwrites the staged inputs
    }
}

// One 16 x 16 tile of pixels per block, so the two reads of a pair's weight,
// by each of its pixels, often fall in one block. The neighbours' colour means
// and variances of the mean (P4.xyz, P3.w) of the tile and its halo are
// copied to LDS first. A window row is walked with its bounds hoisted,
// neighbours before the pixel reading the neighbour's row of the table, those
// after it the pixel's own row, and the weights of four neighbours are loaded
// before any of them is added, in order.
constexpr int STATMC_SUM_TILE = 16;
constexpr int STATMC_SUM_BAND = 8;     // tile rows walked column by column
constexpr int STATMC_SUM_MAX_R = 13;   // the pair form's largest radius (the table must fit the tapes)
constexpr int STATMC_SUM_SPAN = STATMC_SUM_TILE + 2 * STATMC_SUM_MAX_R;
bool statmc_pairs_radius_ok(int r) {
    return r >= 1 && r <= STATMC_SUM_MAX_R &&
           size_t(statmc_forward_count(r)) * sizeof(double) <= size_t(MT_TAPE_WORDS) * sizeof(uint32_t);
}
__global__ void __launch_bounds__(STATMC_SUM_TILE * STATMC_SUM_TILE) k_statmc_sum(Batch b) {
    __shared__ double s_sp[STATMC_SP_MAX + 1];
    __shared__ float4 s_c[STATMC_SUM_SPAN * STATMC_SUM_SPAN];   // neighbour's (P4.xyz, P3.w)
    // The tiles: a job's own image (width1 set: the pair pass may have used
    // another tile shape), or a batch whose BK_TILE row counts 16 x 16 tiles.
    // A pixel reads the weights its neighbours above stored, so a tile's reads
    // follow those of the tile above it. Blocks are dealt to the 8 XCDs in turn
    // (block b on XCD b % 8), so XCD x takes positions x * per .. of one order
    // that goes image by image and, in an image, walks bands of 8 tile rows
    // column by column: the tile above runs one step earlier on the same XCD,
    // whose L2 still holds its weights. The grid is 8 * per blocks.
    const int64_t T = b.width1 > 0 ? int64_t((b.width1 + STATMC_SUM_TILE - 1) / STATMC_SUM_TILE) *
                                         ((b.height1 + STATMC_SUM_TILE - 1) / STATMC_SUM_TILE)
                                   : b.total[BK_TILE];
    const int64_t per = (T + 7) / 8, m = int64_t(blockIdx.x % 8) * per + blockIdx.x / 8;
    if (m >= T) return;
    int64_t lt = m;
    const ImageDesc& d = b.width1 > 0 ? b.img[0] : b.img[batch_find(b, BK_TILE, m, lt)];
    const int W = d.W, H = d.H, n = d.N, R = d.fc.radius;
    const int tiles_x = (W + STATMC_SUM_TILE - 1) / STATMC_SUM_TILE;
    const int tiles_y = (H + STATMC_SUM_TILE - 1) / STATMC_SUM_TILE;
    const int band = int(lt / (int64_t(STATMC_SUM_BAND) * tiles_x));
    const int rows = min(STATMC_SUM_BAND, tiles_y - band * STATMC_SUM_BAND);
    const int64_t within = lt - int64_t(band) * STATMC_SUM_BAND * tiles_x;
    const int tx = int(within / rows), ty = band * STATMC_SUM_BAND + int(within % rows);
    const int x0 = tx * STATMC_SUM_TILE, y0 = ty * STATMC_SUM_TILE;
    const float4* __restrict__ P = reinterpret_cast<const float4*>(d.filter_scratch + 4 * size_t(n));
    const float4* __restrict__ P3 = P + 3 * size_t(n);
    const float4* __restrict__ P4 = P + 4 * size_t(n);
    const int sw = STATMC_SUM_TILE + 2 * R;
    for (int k = threadIdx.x; k <= STATMC_SP_MAX; k += blockDim.x) s_sp[k] = 1.0 / (1.0 + double(k));
    for (int k = threadIdx.x; k < sw * sw; k += blockDim.x) {
        const int gx = x0 - R + k % sw, gy = y0 - R + k / sw;
        if (gx >= 0 && gx < W && gy >= 0 && gy < H) {
            const int q = gy * W + gx;
            const float4 c = P4[q];
            s_c[k] = make_float4(c.x, c.y, c.z, P3[q].w);
        }
    }
    __syncthreads();
    const int x = x0 + int(threadIdx.x) % STATMC_SUM_TILE, y = y0 + int(threadIdx.x) / STATMC_SUM_TILE;
    if (x >= W || y >= H) return;
    const int i = y * W + x;
    const double* __restrict__ Wp = d.pair_w;
    if (P[i].w < 0.0f) {
        statmc_keep_mean(d, i);
        return;
    }
    // the centre's weight first (its inputs are not needed during the walk)
    double w_self;
    {
        const StatmcConst K = statmc_const(d.fc);
        StatmcSelf S;
        statmc_self(d, i, S);
        w_self = statmc_pair_weight(d, K, S, s_sp, i, S.p0, S.p1, 0, 0);
    }
    StatmcSums A;
    const int xa = max(-R, -x), xb = min(R, W - 1 - x);
    const int lx = x - x0 + R;   // the pixel's column in s_c
    for (int dy = -R; dy <= R; ++dy) {
        const int yy = y + dy;
        if (yy < 0 || yy >= H) continue;
        const float4* crow = s_c + (yy - y0 + R) * sw + lx;
        if (dy == 0) {
            for (int dx = xa; dx <= xb; ++dx) {
                double wj;
                if (dx < 0) wj = Wp[size_t(statmc_forward_index(-dx, 0, R)) * n + (i + dx)];
                else if (dx > 0) wj = Wp[size_t(statmc_forward_index(dx, 0, R)) * n + i];
                else wj = w_self;
                if (wj < 0.0) continue;
                const float4 c = crow[dx];
                A.add(wj, statmc_sp(s_sp, dx * dx), c, c.w, dx == 0);
            }
            continue;
        }
        // the neighbour's row (dy < 0: plane of offset (-dx, -dy), one pixel and one plane back per dx)
        // or the pixel's own (dy > 0: plane of offset (dx, dy), one plane on per dx)
        const double* w0 = dy < 0 ? Wp + size_t(statmc_forward_index(-xa, -dy, R)) * n + (i + dy * W + xa)
                                  : Wp + size_t(statmc_forward_index(xa, dy, R)) * n + i;
        const int64_t step = dy < 0 ? 1 - int64_t(n) : int64_t(n);
        const int d2y = dy * dy;
        for (int dx = xa; dx <= xb; dx += 4) {
            double w4[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) w4[u] = dx + u <= xb ? w0[(dx - xa + u) * step] : -1.0;
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                if (w4[u] < 0.0) continue;
                const float4 c = crow[dx + u];
                A.add(w4[u], statmc_sp(s_sp, (dx + u) * (dx + u) + d2y), c, c.w, false);
            }
        }
    }
    StatmcSelf S;
    statmc_self(d, i, S);
    statmc_finish(d, i, S, A);
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
    const void* kern = b.pairs ? reinterpret_cast<const void*>(&k_statmc_pairs<TW, TH>)
                               : reinterpret_cast<const void*>(&k_statmc<TW, TH>);
    static const int cap_one = device_resident_blocks(reinterpret_cast<const void*>(&k_statmc<TW, TH>), TW * TH);
    static const int cap_pairs =
        device_resident_blocks(reinterpret_cast<const void*>(&k_statmc_pairs<TW, TH>), TW * TH);
    const int cap = b.pairs ? cap_pairs : cap_one;
    const size_t lds = b.pairs ? statmc_gate_bytes_forward(TW, TH, b.max_radius) : statmc_gate_bytes(TW, TH, b.max_radius);
    const int stage = b.max_radius > 0 && lds <= STATMC_GATE_LDS_MAX ? 1 : 0;
    int per_cu = 0, dev = 0, cus = 1;
    if (stage && hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, TW * TH, lds) == hipSuccess && per_cu > 0) {
        hipGetDevice(&dev);
        if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0) cus = 1;
    } else {
        per_cu = 0;
    }
    const int64_t limit = stage && per_cu > 0 ? int64_t(per_cu) * cus : cap;
    const int64_t blocks = std::min<int64_t>(b.total[BK_TILE], limit);
    if (blocks <= 0) return;
    if (!b.pairs) {
        hipLaunchKernelGGL((k_statmc<TW, TH>), dim3(unsigned(blocks)), dim3(TW * TH), stage ? lds : 0, st, b, stage);
        return;
    }
    hipLaunchKernelGGL((k_statmc_pairs<TW, TH>), dim3(unsigned(blocks)), dim3(TW * TH), stage ? lds : 0, st, b, stage);
    // the sum pass: one block per 16 x 16 tile (of the job's image, or of the batch)
    const int64_t own_tiles = int64_t((b.width1 + STATMC_SUM_TILE - 1) / STATMC_SUM_TILE) *
                              ((b.height1 + STATMC_SUM_TILE - 1) / STATMC_SUM_TILE);
    const int64_t sum_tiles = ((b.width1 > 0 ? own_tiles : b.total[BK_TILE]) + 7) / 8 * 8;
    hipLaunchKernelGGL(k_statmc_sum, dim3(unsigned(sum_tiles)), dim3(STATMC_SUM_TILE * STATMC_SUM_TILE), 0, st, b);
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
