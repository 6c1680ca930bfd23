// Adaptive sample allocation (path_tracer/main.cpp:1273-1446).
//
// The reference allocates with a sequential FP64 scan: std::accumulate of the
// importance field, an in-order clip-and-redistribute sweep, floors, and a
// systematic residual draw against a running FP64 cumulative sum. Counts are
// reproduced bit for bit here: the two FP64 sums are formed by one lane in
// index order (other lanes stage the data through LDS), clipping is an
// ordered find-first over the current sweep state, and the residual draw is a
// binary search of the sequential prefix sums, which is exactly where the
// reference's forward cursor stops.
#include "kernels.h"

namespace rpd {

// main.cpp:1305-1328: variance source and sqrt importance
__global__ void k_importance(int n, PixelBuffers pb, const float* vden, float smax, float* imp) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int cnt = pb.n[i];
    float raw = cnt > 1 ? pb.lum_m2[i] / float(cnt - 1) : 0.0f;
    if (!isfinite(raw)) raw = RPD_EPS_SMALL;
    raw = fmaxf(RPD_EPS_SMALL, raw);
    float d = vden[i];
    if (!isfinite(d) || d <= 0.0f) d = raw;
    imp[i] = sqrtf(clampf(fmaxf(raw, d), RPD_EPS_SMALL, smax * smax));
}

// main.cpp:1331-1368: edge-aware smoothing
__global__ void k_importance_smooth(int W, int H, int R, PixelBuffers pb, Features f, const float* src, float* dst) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int n = W * H;
    if (i >= n) return;
    if (pb.hit_count[i] == 0) { dst[i] = src[i]; return; }
    const int x = i % W, y = i / W;
    const v3 ni = mk3(f.normal[i], f.normal[n + i], f.normal[2 * n + i]);
    const float di = f.depth[i];
    double acc = 0.0, ws = 0.0;
    for (int dy = -R; dy <= R; ++dy) {
        const int yy = y + dy;
        if (yy < 0 || yy >= H) continue;
        for (int dx = -R; dx <= R; ++dx) {
            const int xx = x + dx;
            if (xx < 0 || xx >= W) continue;
            const int q = yy * W + xx;
            if (pb.hit_count[q] == 0) continue;
            const float sw = 1.0f / (1.0f + float(dx * dx + dy * dy));
            const float nw = clampf((dot3(ni, mk3(f.normal[q], f.normal[n + q], f.normal[2 * n + q])) + 1.0f) * 0.5f,
                                    0.0f, 1.0f);
            const float dw = expf(-4.0f * relative_depth_delta(di, f.depth[q]));
            const double w = double(sw) * double(nw) * double(dw);
            acc += w * double(src[q]);
            ws += w;
        }
    }
    dst[i] = (ws > double(RPD_EPS_SMALL)) ? float(acc / ws) : src[i];
}

constexpr int ALLOC_BLOCK = 1024;
constexpr int STAGE = 4096;  // floats staged per LDS buffer

// Sequential FP64 sum (and optionally the running prefix) of a float array in
// index order. Thread 0 adds; the other threads stage the next chunk into the
// second LDS buffer and write the previous chunk's prefix values out.
__device__ double seq_sum(const float* __restrict__ v, int n, double* prefix, float* stage /* 2*STAGE */,
                          double* pstage /* STAGE */) {
    const int t = threadIdx.x;
    double s = 0.0;
    const int nchunks = (n + STAGE - 1) / STAGE;
    for (int k = t; k < STAGE; k += ALLOC_BLOCK) stage[k] = k < n ? v[k] : 0.0f;
    __syncthreads();
    for (int c = 0; c < nchunks; ++c) {
        const float* cur = stage + (c & 1) * STAGE;
        float* nxt = stage + ((c + 1) & 1) * STAGE;
        const int base = c * STAGE;
        const int cnt = min(STAGE, n - base);
        if (t == 0) {
            if (prefix) {
                for (int k = 0; k < cnt; ++k) { s += double(cur[k]); pstage[k] = s; }
            } else {
                int k = 0;
                for (; k + 4 <= cnt; k += 4) {
                    const float a = cur[k], b = cur[k + 1], cc = cur[k + 2], d = cur[k + 3];
                    s += double(a); s += double(b); s += double(cc); s += double(d);
                }
                for (; k < cnt; ++k) s += double(cur[k]);
            }
        } else if (c + 1 < nchunks) {
            const int nb = base + STAGE;
            for (int k = t - 1; k < STAGE; k += ALLOC_BLOCK - 1) nxt[k] = nb + k < n ? v[nb + k] : 0.0f;
        }
        __syncthreads();
        if (prefix) {
            for (int k = t; k < cnt; k += ALLOC_BLOCK) prefix[base + k] = pstage[k];
            __syncthreads();
        }
    }
    return s;  // valid in thread 0
}

struct AllocScratch {
    float* desired;
    unsigned char* clipped;
    double* prefix;
    int* misc;  // [0] status
};

__global__ void __launch_bounds__(ALLOC_BLOCK) k_alloc_serial(int n, AllocConfig ac, const float* imp, int* counts,
                                                             AllocScratch s) {
    __shared__ float stage[2 * STAGE];
    __shared__ double pstage[STAGE];
    __shared__ double sh_d[4];
    __shared__ int sh_i[4];
    __shared__ int wave_first[ALLOC_BLOCK / 64];
    const int t = threadIdx.x;
    const int total = ac.adaptive_spp * n;
    const int floor_pp = min(max(ac.base_samples, 0), ac.adaptive_spp);
    const int reserved = floor_pp * n;
    for (int i = t; i < n; i += ALLOC_BLOCK) { counts[i] = floor_pp; s.desired[i] = 0.0f; s.clipped[i] = 0; }
    if (total <= 0) return;
    const double sum = seq_sum(imp, n, nullptr, stage, pstage);
    if (t == 0) sh_d[0] = sum;
    __syncthreads();
    double rem_imp = sh_d[0];
    int remaining = max(0, total - reserved);
    if (rem_imp <= 0.0 || remaining <= 0) return;
    const int max_pp = max(4, 8 * max(1, ac.adaptive_spp));

    // clip-and-redistribute sweeps (main.cpp:1387-1412): every sweep visits the
    // unclipped pixels in index order; a pixel whose share of the remaining
    // samples exceeds the cap is clipped and the state it changes applies to
    // the rest of the sweep, so each clip is the first index at or after the
    // previous one that satisfies the test under the current state.
    int active_count = n;
    while (active_count > 0 && remaining > 0 && rem_imp > double(RPD_EPS_SMALL)) {
        bool clipped_any = false;
        int pos = 0;
        while (pos < n) {
            int found = 0x7fffffff;
            for (int base = pos; base < n && found == 0x7fffffff; base += ALLOC_BLOCK) {
                const int i = base + t;
                bool hitc = false;
                if (i < n && !s.clipped[i]) {
                    const double want = double(imp[i]) / rem_imp * double(remaining);
                    hitc = want > double(max_pp);
                }
                const uint64_t b = __ballot(hitc);
                if ((t & 63) == 0) wave_first[t >> 6] = b ? (base + (t & ~63) + __ffsll((long long)b) - 1) : 0x7fffffff;
                __syncthreads();
                if (t == 0) {
                    int f = 0x7fffffff;
                    for (int w = 0; w < ALLOC_BLOCK / 64; ++w) f = min(f, wave_first[w]);
                    sh_i[0] = f;
                }
                __syncthreads();
                found = sh_i[0];
                __syncthreads();
            }
            if (found == 0x7fffffff) break;
            if (t == 0) {
                s.desired[found] = float(max_pp);
                s.clipped[found] = 1;
            }
            remaining -= max_pp;
            rem_imp -= double(imp[found]);
            --active_count;
            clipped_any = true;
            pos = found + 1;
            __syncthreads();
        }
        if (!clipped_any) {
            for (int i = t; i < n; i += ALLOC_BLOCK)
                if (!s.clipped[i]) s.desired[i] = float(double(imp[i]) / rem_imp * double(remaining));
            break;
        }
    }
    __syncthreads();

    // floors and residuals (main.cpp:1414-1423)
    int assigned_local = 0;
    for (int i = t; i < n; i += ALLOC_BLOCK) {
        const float d = s.desired[i];
        const int add = int(floorf(d));
        counts[i] += add;
        assigned_local += add;
        s.desired[i] = fmaxf(0.0f, d - float(add));  // residual weight
    }
    __shared__ int red[ALLOC_BLOCK];
    red[t] = assigned_local;
    __syncthreads();
    for (int o = ALLOC_BLOCK / 2; o > 0; o >>= 1) {
        if (t < o) red[t] += red[t + o];
        __syncthreads();
    }
    const int assigned = red[0];
    __syncthreads();
    const double rsum_t0 = seq_sum(s.desired, n, s.prefix, stage, pstage);
    if (t == 0) sh_d[1] = rsum_t0;
    __syncthreads();
    const double rsum = sh_d[1];
    const int left = max(0, total - reserved - assigned);
    if (left <= 0 || !(rsum > double(RPD_EPS_SMALL))) return;
    // systematic residual sampling (main.cpp:1425-1438)
    const double off = double(hash_to_unit_float(ac.seed ^ 0x4d595df4u));
    for (int k = t; k < left; k += ALLOC_BLOCK) {
        const double target = ((double(k) + off) / double(left)) * rsum;
        int lo = 0, hi = n;  // first c with prefix[c] >= target
        while (lo < hi) { const int mid = (lo + hi) >> 1; if (s.prefix[mid] < target) lo = mid + 1; else hi = mid; }
        if (lo < n) atomicAdd(&counts[lo], 1);
    }
}

size_t alloc_scratch_bytes(int n) {
    size_t b = 0;
    b += ((size_t)n * sizeof(float) + 255) & ~size_t(255);
    b += ((size_t)n + 255) & ~size_t(255);
    b += ((size_t)n * sizeof(double) + 255) & ~size_t(255);
    b += 256;
    b += ((size_t)n * sizeof(float) + 255) & ~size_t(255);  // smoothing buffer
    return b;
}

void launch_allocation(int W, int H, const AllocConfig& ac, const PixelBuffers& pb, const Features& f,
                       const float* vden, float* imp, int* counts, void* scratch, hipStream_t st) {
    const int n = W * H;
    char* p = static_cast<char*>(scratch);
    AllocScratch s;
    s.desired = reinterpret_cast<float*>(p); p += ((size_t)n * sizeof(float) + 255) & ~size_t(255);
    s.clipped = reinterpret_cast<unsigned char*>(p); p += ((size_t)n + 255) & ~size_t(255);
    s.prefix = reinterpret_cast<double*>(p); p += ((size_t)n * sizeof(double) + 255) & ~size_t(255);
    s.misc = reinterpret_cast<int*>(p); p += 256;
    float* tmp = reinterpret_cast<float*>(p);
    const dim3 grid((n + 255) / 256);
    if (ac.adaptive_spp > 0) {
        hipLaunchKernelGGL(k_importance, grid, dim3(256), 0, st, n, pb, vden, ac.sigma_max, imp);
        if (ac.radius > 0) {
            hipLaunchKernelGGL(k_importance_smooth, grid, dim3(256), 0, st, W, H, ac.radius, pb, f, imp, tmp);
            hipMemcpyAsync(imp, tmp, (size_t)n * sizeof(float), hipMemcpyDeviceToDevice, st);
        }
    }
    hipLaunchKernelGGL(k_alloc_serial, dim3(1), dim3(ALLOC_BLOCK), 0, st, n, ac, imp, counts, s);
}

}  // namespace rpd
