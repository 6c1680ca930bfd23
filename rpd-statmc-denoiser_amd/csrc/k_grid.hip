// RPF dependency grid and sensitivity reconstruction.
//
// Reference: path_tracer/main.cpp:329-388 (splatting samples into the
// overlapping grid), core/stats.h:93-329 (binned dependency statistics),
// main.cpp:741-953 (recompute_sensitivity), core/buffers.h:68-101 (bilinear
// node -> pixel expansion), main.cpp:94-122 (feature averages).
#include "kernels.h"

#include <algorithm>

namespace rpd {

// ---------------------------------------------------------------------------
// Per-pixel feature averages: main.cpp:106-118
__global__ void k_features(Batch b) {
    const int64_t item = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (item >= b.total[BK_PIXEL]) return;
    int64_t li;
    const ImageDesc& d = b.img[batch_find(b, BK_PIXEL, item, li)];
    const int i = int(li), n = d.N;
    const PixelBuffers& pb = d.pb;
    const Features& f = d.feat;
    const int hc = pb.hit_count[i];
    v3 nr = mk3(pb.normal[i], pb.normal[n + i], pb.normal[2 * n + i]);
    v3 al = mk3(pb.albedo[i], pb.albedo[n + i], pb.albedo[2 * n + i]);
    float dd = pb.depth[i];
    if (hc > 0) {
        const float h = float(hc);
        nr = nr / h;
        if (sqnorm3(nr) > RPD_EPS_SMALL) nr = normalized3(nr);
        dd = dd / h;
        al = al / h;
    }
    f.normal[i] = nr.x; f.normal[n + i] = nr.y; f.normal[2 * n + i] = nr.z;
    f.depth[i] = dd;
    f.albedo[i] = al.x; f.albedo[n + i] = al.y; f.albedo[2 * n + i] = al.z;
}
void launch_features(const Batch& b, hipStream_t st) {
    hipLaunchKernelGGL(k_features, dim3((b.total[BK_PIXEL] + 255) / 256), dim3(256), 0, st, b);
}

// ---------------------------------------------------------------------------
__global__ void k_queues_zero(unsigned int* q) {
    if (threadIdx.x < BK_KINDS * BATCH_QUEUES) q[threadIdx.x] = 0u;
}
__global__ void k_set_prefix_row(int64_t* row, int64_t total) {
    if (threadIdx.x == 0) { row[0] = 0; row[1] = total; }
}
void launch_set_tile_prefix1(const Batch& b, int64_t tiles, hipStream_t st) {
    hipLaunchKernelGGL(k_set_prefix_row, dim3(1), dim3(64), 0, st, const_cast<int64_t*>(b.prefix) + BK_TILE * 2,
                       tiles);
}

// A batch of one's device copies, written on the stream: the image's
// descriptor (passed by value in the kernel arguments), its prefix rows and
// zeroed work counters. No host round trip, unlike a synchronous copy.
__global__ void k_put_batch1(ImageDesc* dd, ImageDesc d, int64_t* dp, BatchPrefix1 pf, unsigned int* dq) {
    if (threadIdx.x == 0) *dd = d;
    if (threadIdx.x < BK_KINDS * 2) dp[threadIdx.x] = pf.v[threadIdx.x];
    if (threadIdx.x < BK_KINDS * BATCH_QUEUES) dq[threadIdx.x] = 0u;
}
void launch_put_batch1(ImageDesc* dd, const ImageDesc& d, int64_t* dp, const BatchPrefix1& pf, unsigned int* dq,
                       hipStream_t st) {
    hipLaunchKernelGGL(k_put_batch1, dim3(1), dim3(64), 0, st, dd, d, dp, pf, dq);
}

void launch_batch_queues_reset(const Batch& b, hipStream_t st) {
    hipLaunchKernelGGL(k_queues_zero, dim3(1), dim3(64), 0, st, b.queues);
}

// ---------------------------------------------------------------------------
// Grid gather. One workgroup owns one node and walks the pixels of the node's
// support in row-major order and each pixel's samples in order, which is the
// order a single sampling thread of the reference feeds that node. Phase 1
// turns up to 256 (pixel, sample) contributions into their FP64 terms (w,
// w*y, w*w, w*y*y) and 13 bin indices in LDS; phase 2 gives each (accumulator,
// bin) pair and each accumulator's two total pairs one thread that adds the
// chunk sequentially in FP64, so every sum is formed in sample order without
// atomics. The pass totals are then added to the node.
//
// The first threads hold the bin sums of the accumulators the image's samples
// can reach, the next ones their totals (at most 220 threads for all 13
// accumulators). Every such thread runs the same loop: it reads its own pair
// of terms ((w, w*y) for a bin, (w, w*w) or (w*y, w*y*y) for the totals),
// compares the contribution's bin with its key (its own bin, or -1, "no
// bin", for the totals) and picks m from two per-thread values by the result
// (bins: equal gives 1; totals: equal gives -0), and adds with
// fma(m, term, sum), m = 1 for a member (fma(1, x, s) rounds x + s once,
// exactly the addition) and m = -0 otherwise (a sum that starts at +0 never
// becomes -0 under round-to-nearest, so adding the zero -0 * term leaves it
// unchanged). That needs finite terms, and they are: weights lie in [0, 1],
// and the luminance of finite channels is finite, because its float weights
// sum to exactly 1 and its unfused evaluation is monotone in each channel, so
// |y| <= luminance3(FLT_MAX, FLT_MAX, FLT_MAX) = FLT_MAX. The support list is
// built by all threads at once (weights, counts, then an ordered block scan).
constexpr int GATHER_BLOCK = 256;
// The support window of a node holds up to (2*step)^2 pixels. The kernel is
// built for windows of 64, 256 and 1024 pixels (steps up to 4, 8 and 16) and
// launched with the smallest that holds every image of the batch, so the
// default step of 4 leaves LDS room for more resident blocks; a larger step
// walks its window in segments of 1024 positions.
constexpr int GATHER_MAX_STEP = 16;
__host__ __device__ constexpr int gather_max_pix(int max_step) { return 4 * max_step * max_step; }

__device__ __forceinline__ int dep_bin(float u, float v) {  // core/stats.h:165-169
    const int x = min(max(int(u * 4), 0), 3);
    const int z = min(max(int(v * 4), 0), 3);
    return z * 4 + x;
}

// Exclusive block scan of one 64-bit value per thread (GATHER_BLOCK threads):
// each wave scans its 64 values with lane shuffles, the wave totals meet in
// LDS, and every thread adds those of the waves before it. `total` receives
// the block's sum. Two barriers; `sh` holds one entry per wave.
__device__ __forceinline__ int64_t gather_scan(int64_t v, int64_t* sh, int64_t& total) {
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    int64_t x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int64_t y = __shfl_up(x, o);
        if (lane >= o) x += y;
    }
    if (lane == 63) sh[wave] = x;
    __syncthreads();
    int64_t before = 0;
    total = 0;
#pragma unroll
    for (int w = 0; w < GATHER_BLOCK / 64; ++w) {
        const int64_t tw = sh[w];
        if (w < wave) before += tw;
        total += tw;
    }
    __syncthreads();    // sh is reused by the next scan
    return before + x - v;
}

constexpr int GATHER_BIN_ROW = GATHER_BLOCK / 4 + 1;  // dwords per accumulator row of bin bytes

template <int kMaxPix>
struct GatherLds {
    int pix[kMaxPix];
    float w[kMaxPix];
    int pre[kMaxPix + 1];
    int64_t scan[GATHER_BLOCK / 64];
    // Rows are padded by one element so that the rows one wave reads at the
    // same index (several accumulators' bins, the three pair arrays) start in
    // different LDS banks.
    double2 pr[3][GATHER_BLOCK + 1];  // per contribution: (w, w*y), (w, w*w), (w*y, w*y*y)
    uint32_t bin4[GRID_ACC][GATHER_BIN_ROW];
    int64_t item;
};
// every accumulator's bins and total pairs have a thread of the block
static_assert((GRID_ACC - 1) * 16 + grid_acc_bins(11) + 2 * GRID_ACC <= GATHER_BLOCK, "gather sum threads");

// One node of one image; called by all threads of the block.
template <int kMaxPix>
__device__ void gather_node(const ImageDesc& d, const int node, GatherLds<kMaxPix>& L) {
    const GridGeom g = d.g;
    const int W = d.W, H = d.H;
    const PassPlan& plan = d.plan;
    const Records& rec = d.rec;
    double* grid = d.grid;
    int* s_pix = L.pix;
    float* s_w = L.w;
    int* s_pre = L.pre;
    int64_t* s_scan = L.scan;
    signed char* c_bin = reinterpret_cast<signed char*>(L.bin4);

    const int gx = node % g.nx, gy = node / g.nx;
    const int tid = threadIdx.x;
    const int x_lo = max(0, (gx - 1) * g.step), x_hi = min(W - 1, (gx + 1) * g.step - 1);
    const int y_lo = max(0, (gy - 1) * g.step), y_hi = min(H - 1, (gy + 1) * g.step - 1);
    const int wx = x_hi - x_lo + 1, wy = y_hi - y_lo + 1;
    const int nwin = (wx > 0 && wy > 0) ? wx * wy : 0;
    const bool any_active = plan.totals[1] > 0;

    // This thread's sum: a bin (acc, bsel < 16) or a total pair of accumulator
    // acc (bsel 16: w and w*w, bsel 17: w*y and w*y*y). Only the accumulators
    // the image's samples can reach (acc_mask; the environment's when the
    // scene has an environment, the lens's when the camera has an aperture)
    // get threads: first every such accumulator's bins, then their totals, so
    // that the threads past the last sum form whole waves that skip the sums.
    // An accumulator no sample reaches keeps the zeros it starts with.
    int acc = -1, bsel = 0;
    {
        int t = tid;
        for (int k = 0; k < GRID_ACC && acc < 0; ++k) {
            if (!((d.acc_mask >> k) & 1u)) continue;
            if (t < grid_acc_bins(k)) { acc = k; bsel = t; }
            else t -= grid_acc_bins(k);
        }
        for (int k = 0; k < GRID_ACC && acc < 0; ++k) {
            if (!((d.acc_mask >> k) & 1u)) continue;
            if (t < 2) { acc = k; bsel = 16 + t; }
            else t -= 2;
        }
    }
    const bool active_t = acc >= 0;
    const bool bin_t = bsel < 16;
    double a0 = 0.0, a1 = 0.0;

    const float half = 0.5f * float(g.support);
    const float left = float(gx * g.step) - half, top = float(gy * g.step) - half;

    // The window is walked in segments of at most kMaxPix positions (one
    // segment unless the grid step exceeds the kernel's window), each with its
    // own support list; the sums run on across segments, so the order of the
    // contributions is the same.
    int grand = 0;
    for (int seg = 0; seg < nwin; seg += kMaxPix) {
        const int seg_end = min(nwin, seg + kMaxPix);
        // support pixels with their node weight (main.cpp:332-349), row-major,
        // over window positions [seg, seg_end)
        int npix = 0, total = 0;
        for (int q0 = seg; q0 < seg_end; q0 += GATHER_BLOCK) {
            const int q = q0 + tid;
            float w = 0.0f;
            int c = 0, p = 0;
            if (q < seg_end) {
                const int x = x_lo + q % wx, y = y_lo + q / wx;
                const float fx = float(x) / float(g.step), fy = float(y) / float(g.step);
                const int x0 = min(max(int(floorf(fx)), 0), g.nx - 1), y0 = min(max(int(floorf(fy)), 0), g.ny - 1);
                const int x1 = min(x0 + 1, g.nx - 1), y1 = min(y0 + 1, g.ny - 1);
                const float tx = clampf(fx - float(x0), 0.0f, 1.0f), ty = clampf(fy - float(y0), 0.0f, 1.0f);
                if (x0 == gx && y0 == gy) w = (1.0f - tx) * (1.0f - ty);
                else if (x1 == gx && y0 == gy) w = tx * (1.0f - ty);
                else if (x0 == gx && y1 == gy) w = (1.0f - tx) * ty;
                else if (x1 == gx && y1 == gy) w = tx * ty;
                p = y * W + x;
                if (w > RPD_EPS_SMALL && any_active) c = max(plan.counts[p], 0);
            }
            const bool keep = c > 0;
            // kept pixels (high word) and their sample counts (low word) scanned together
            int64_t both = 0;
            const int64_t pos = gather_scan((int64_t(keep ? 1 : 0) << 32) | int64_t(c), s_scan, both);
            const int kpos = int(pos >> 32), cpos = int(pos & 0xffffffff);
            if (keep) { s_pix[npix + kpos] = p; s_w[npix + kpos] = w; s_pre[npix + kpos] = total + cpos; }
            npix += int(both >> 32);
            total += int(both & 0xffffffff);
        }
        if (tid == 0) s_pre[npix] = total;
        __syncthreads();

        for (int base = 0; base < total; base += GATHER_BLOCK) {
            const int c = base + tid;
            signed char bins[GRID_ACC];
            for (int k = 0; k < GRID_ACC; ++k) bins[k] = -1;
            float y = 0.0f, wc = 0.0f;
            if (c < total) {
                int lo = 0, hi = npix - 1;  // pixel of contribution c: last j with s_pre[j] <= c
                while (lo < hi) { const int mid = (lo + hi + 1) >> 1; if (s_pre[mid] <= c) lo = mid; else hi = mid - 1; }
                const int j = lo;
                const int p = s_pix[j];
                wc = s_w[j];
                const int64_t slot = int64_t(plan.offsets[p]) + (c - s_pre[j]);
                // all four planes at once: one memory round trip (the last three
                // are unused for a non-finite sample)
                const float4 r0 = rec.r0[slot];
                const float4 r4 = rec.r4[slot];
                const float4 r5 = rec.r5[slot];
                const float2 r6 = rec.r6[slot];
                const uint32_t fl = __float_as_uint(r0.w);
                if (fl & REC_FINITE) {
                    y = luminance3(mk3(r0.x, r0.y, r0.z));
                    const int x = p % W, yy = p / W;
                    if (fl & REC_PIXEL) {
                        const float su = clampf((float(x) + r4.x - left) / float(g.support), 0.0f, 1.0f);
                        const float sv = clampf((float(yy) + r4.y - top) / float(g.support), 0.0f, 1.0f);
                        const int sb = dep_bin(su, sv);
                        bins[0] = (signed char)sb;
                        if (fl & REC_BRDF) bins[2] = (signed char)sb;
                        if (fl & REC_LENS) bins[4] = (signed char)sb;
                        if (fl & REC_LIGHT) bins[6] = (signed char)sb;
                        if (fl & REC_ENV) bins[8] = (signed char)sb;
                        if (fl & REC_LSEL) bins[10] = (signed char)sb;
                        if (fl & REC_RR) bins[12] = (signed char)sb;
                    }
                    if (fl & REC_BRDF) bins[1] = (signed char)dep_bin(r5.x, r5.y);
                    if (fl & REC_LENS) bins[3] = (signed char)dep_bin(r4.z, r4.w);
                    if (fl & REC_LIGHT) bins[5] = (signed char)dep_bin(r5.z, r5.w);
                    if (fl & REC_ENV) bins[7] = (signed char)dep_bin(r6.x, r6.y);
                    if (fl & REC_LSEL) bins[9] = (signed char)min(max(int((fl >> REC_LIGHT_SHIFT) & 255u), 0), 15);
                    if (fl & REC_RR) bins[11] = (signed char)((fl & REC_RR_SURVIVED) ? 1 : 0);
                }
            }
            // entries past the end of the chunk carry bin -1 and add nothing; the
            // FP64 terms are the products the sums add, each rounded once
            {
                const double wd = double(wc), yd = double(y);
                const double wyd = wd * yd;
                L.pr[0][tid] = make_double2(wd, wyd);
                L.pr[1][tid] = make_double2(wd, wd * wd);
                L.pr[2][tid] = make_double2(wyd, wyd * yd);
            }
            for (int k = 0; k < GRID_ACC; ++k) c_bin[k * 4 * GATHER_BIN_ROW + tid] = bins[k];
            __syncthreads();
            const int nquad = (min(GATHER_BLOCK, total - base) + 3) >> 2;
            if (active_t) {
                const double2* src = L.pr[bin_t ? 0 : bsel - 15];
                const int key = bin_t ? bsel : -1;
                const double m_eq = bin_t ? 1.0 : -0.0, m_ne = bin_t ? -0.0 : 1.0;
                for (int i4 = 0; i4 < nquad; ++i4) {
                    const uint32_t b4 = L.bin4[acc][i4];
#pragma unroll
                    for (int e = 0; e < 4; ++e) {
                        const int bin = int(int8_t(b4 >> (8 * e)));  // -1: none
                        const double m = bin == key ? m_eq : m_ne;
                        const double2 v = src[4 * i4 + e];
                        a0 = fma(m, v.x, a0);
                        a1 = fma(m, v.y, a1);
                    }
                }
            }
            __syncthreads();
        }
        grand += total;
    }
    if (active_t && grand > 0) {
        double* nd = grid + (size_t)node * GRID_NODE_DOUBLES + grid_acc_offset(acc);
        const int nb = grid_acc_bins(acc);
        if (bsel < 16) { nd[bsel] += a0; nd[nb + bsel] += a1; }
        else if (bsel == 16) { nd[2 * nb] += a0; nd[2 * nb + 1] += a1; }
        else { nd[2 * nb + 2] += a0; nd[2 * nb + 3] += a1; }
    }
}

template <int kMaxStep>
__global__ void __launch_bounds__(GATHER_BLOCK, 8) k_grid_gather(Batch b) {
    __shared__ GatherLds<gather_max_pix(kMaxStep)> L;
    int cursor = 0;
    while (true) {
        if (threadIdx.x == 0) L.item = batch_take(b, BK_NODE, cursor);
        __syncthreads();
        const int64_t item = L.item;
        __syncthreads();
        if (item < 0) break;
        int64_t node;
        const ImageDesc& d = b.img[batch_find(b, BK_NODE, item, node)];
        gather_node(d, int(node), L);
    }
}

int device_resident_blocks(const void* kernel, int block_threads) {
    int per_cu = 0, cus = 0, dev = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, block_threads, 0) != hipSuccess || per_cu <= 0)
        per_cu = 1;
    hipGetDevice(&dev);
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0) cus = 1;
    return per_cu * cus;
}

template <int kMaxStep>
static void launch_grid_gather_t(const Batch& b, hipStream_t st) {
    // resident blocks over the device; a function-local static is initialised
    // once even when several job threads launch at the same time
    static const int cap = device_resident_blocks(reinterpret_cast<const void*>(&k_grid_gather<kMaxStep>), GATHER_BLOCK);
    const int64_t blocks = std::min<int64_t>(b.total[BK_NODE], cap);
    if (blocks <= 0) return;
    hipLaunchKernelGGL(k_grid_gather<kMaxStep>, dim3(unsigned(blocks)), dim3(GATHER_BLOCK), 0, st, b);
}

// b.max_step <= 0 (not known) takes the largest window
void launch_grid_gather(const Batch& b, hipStream_t st) {
    if (b.max_step > 0 && b.max_step <= 4) launch_grid_gather_t<4>(b, st);
    else if (b.max_step > 0 && b.max_step <= 8) launch_grid_gather_t<8>(b, st);
    else launch_grid_gather_t<GATHER_MAX_STEP>(b, st);
}

// ---------------------------------------------------------------------------
// core/stats.h:115-137: bias-adjusted explained variance of one accumulator
__device__ inline float binned_sensitivity(const double* a, int nb) {
    const double kEps = 1e-18;
    const double ws = a[2 * nb], w2 = a[2 * nb + 1], ys = a[2 * nb + 2], y2 = a[2 * nb + 3];
    const double ne = w2 > 0.0 ? ws * ws / w2 : 0.0;
    if (ne <= 2.0) return 0.0f;
    const double mean = ys / ws;
    const double tot = y2 - ws * mean * mean;
    if (tot <= kEps) return 0.0f;
    int active = 0;
    double between = 0.0;
    for (int b = 0; b < nb; ++b) {
        const double w = a[b];
        if (w <= kEps) continue;
        ++active;
        const double d = a[nb + b] / w - mean;
        between += w * d * d;
    }
    if (active < 2 || ne <= double(active)) return 0.0f;
    const double eta2 = fmin(fmax(between / tot, 0.0), 1.0);
    const double adj = fmax(0.0, 1.0 - (1.0 - eta2) * (ne - 1.0) / (ne - active));
    return float(sqrt(adj));
}
__device__ inline double binned_mass(const double* a, int nb) {
    const double ws = a[2 * nb], w2 = a[2 * nb + 1];
    return w2 > 0.0 ? ws * ws / w2 : 0.0;
}

// 3x3 clamped local averages with weights 1/(1+d^2) (main.cpp:156-185)
__device__ inline float local_avg1(const float* v, int w, int h, int cx, int cy) {
    double acc = 0.0, ws = 0.0;
    for (int dy = -1; dy <= 1; ++dy) {
        const int y = min(max(cy + dy, 0), h - 1);
        for (int dx = -1; dx <= 1; ++dx) {
            const int x = min(max(cx + dx, 0), w - 1);
            const double wt = 1.0 / (1.0 + double(dx * dx + dy * dy));
            acc += wt * double(v[y * w + x]);
            ws += wt;
        }
    }
    return ws > 0.0 ? float(acc / ws) : 0.0f;
}
__device__ inline v3 local_avg3(const float* v, int n, int w, int h, int cx, int cy) {
    double a0 = 0, a1 = 0, a2 = 0, ws = 0;
    for (int dy = -1; dy <= 1; ++dy) {
        const int y = min(max(cy + dy, 0), h - 1);
        for (int dx = -1; dx <= 1; ++dx) {
            const int x = min(max(cx + dx, 0), w - 1);
            const double wt = 1.0 / (1.0 + double(dx * dx + dy * dy));
            const int q = y * w + x;
            a0 += double(v[q]) * wt; a1 += double(v[n + q]) * wt; a2 += double(v[2 * n + q]) * wt;
            ws += wt;
        }
    }
    if (ws <= 0.0) return mk3(0, 0, 0);
    return mk3(float(a0 / ws), float(a1 / ws), float(a2 / ws));
}

// node scratch planes: raw[13], conf[13], normal[3], depth, albedo[3]
constexpr int NODE_PLANES = 33;

// main.cpp:773-815
__global__ void k_node_split(Batch b) {
    const int64_t item = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (item >= b.total[BK_NODE]) return;
    int64_t li;
    const ImageDesc& dsc = b.img[batch_find(b, BK_NODE, item, li)];
    const GridGeom g = dsc.g;
    const int W = dsc.W, H = dsc.H;
    const float n0 = dsc.conf_n0;
    const double* grid = dsc.grid;
    const Features& f = dsc.feat;
    float* ns = dsc.node_scratch;
    const int idx = int(li);
    const int nn = g.nx * g.ny;
    const double* nd = grid + (size_t)idx * GRID_NODE_DOUBLES;
    const double n0d = double(n0);
    for (int k = 0; k < GRID_ACC; ++k) {
        const double* a = nd + grid_acc_offset(k);
        const int nb = grid_acc_bins(k);
        ns[k * nn + idx] = binned_sensitivity(a, nb);
        const double m = binned_mass(a, nb);
        ns[(GRID_ACC + k) * nn + idx] = float(m / (m + n0d));
    }
    const int gx = idx % g.nx, gy = idx / g.nx;
    const int px = min(max(gx * g.step, 0), W - 1), py = min(max(gy * g.step, 0), H - 1);
    const int np = W * H;
    v3 nr = local_avg3(f.normal, np, W, H, px, py);
    if (sqnorm3(nr) > RPD_EPS_SMALL) nr = normalized3(nr);
    const float d = local_avg1(f.depth, W, H, px, py);
    const v3 al = local_avg3(f.albedo, np, W, H, px, py);
    ns[26 * nn + idx] = nr.x; ns[27 * nn + idx] = nr.y; ns[28 * nn + idx] = nr.z;
    ns[29 * nn + idx] = d;
    ns[30 * nn + idx] = al.x; ns[31 * nn + idx] = al.y; ns[32 * nn + idx] = al.z;
}

// main.cpp:817-928: confidence-weighted 3x3 prior per field, shrinkage toward
// it, then reliable random components relative to their screen comparators.
__global__ void k_prior(Batch b) {
    const int64_t item = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (item >= b.total[BK_NODE]) return;
    int64_t li;
    const ImageDesc& dsc = b.img[batch_find(b, BK_NODE, item, li)];
    const GridGeom g = dsc.g;
    const float* ns = dsc.node_scratch;
    float* tiles = dsc.tiles;
    const int idx = int(li);
    const int nn = g.nx * g.ny;
    const int gx = idx % g.nx, gy = idx / g.nx;
    const v3 ni = mk3(ns[26 * nn + idx], ns[27 * nn + idx], ns[28 * nn + idx]);
    const float di = ns[29 * nn + idx];
    const v3 ai = mk3(ns[30 * nn + idx], ns[31 * nn + idx], ns[32 * nn + idx]);
    // neighbour weights without the confidence factor are shared by all fields
    double wgt[9];
    int nidx[9];
    for (int dy = -1; dy <= 1; ++dy) {
        for (int dx = -1; dx <= 1; ++dx) {
            const int q = (dy + 1) * 3 + (dx + 1);
            const int sx = gx + dx, sy = gy + dy;
            if (sx < 0 || sx >= g.nx || sy < 0 || sy >= g.ny) { nidx[q] = -1; wgt[q] = 0.0; continue; }
            const int si = sy * g.nx + sx;
            const v3 nj = mk3(ns[26 * nn + si], ns[27 * nn + si], ns[28 * nn + si]);
            const v3 aj = mk3(ns[30 * nn + si], ns[31 * nn + si], ns[32 * nn + si]);
            const float sw = 1.0f / (1.0f + float(dx * dx + dy * dy));
            const float nw = clampf((dot3(ni, nj) + 1.0f) * 0.5f, 0.0f, 1.0f);
            const float dw = expf_cr(-4.0f * relative_depth_delta(di, ns[29 * nn + si]));
            const float aw = expf_cr(-2.5f * norm3(ai - aj));
            nidx[q] = si;
            wgt[q] = double(sw) * double(nw) * double(dw) * double(aw);
        }
    }
    float fin[GRID_ACC];
    for (int k = 0; k < GRID_ACC; ++k) {
        const float* raw = ns + k * nn;
        const float* cf = ns + (GRID_ACC + k) * nn;
        double ws = 0.0, acc = 0.0;
        for (int q = 0; q < 9; ++q) {
            if (nidx[q] < 0) continue;
            const double w = wgt[q] * double(fmaxf(0.05f, cf[nidx[q]]));
            acc += w * double(raw[nidx[q]]);
            ws += w;
        }
        const float prior = (ws > double(RPD_EPS_SMALL)) ? float(acc / ws) : raw[idx];
        fin[k] = lerpf_(prior, raw[idx], cf[idx]);
    }
    // accumulator indices: 0 pixel, 1 brdf, 2 s_brdf, 3 lens, 4 s_lens, 5 light_uv, 6 s_light_uv,
    // 7 env, 8 s_env, 9 light_select, 10 s_light_select, 11 rr, 12 s_rr
    const int rnd[6] = {1, 3, 5, 9, 7, 11};
    const int scr[6] = {2, 4, 6, 10, 8, 12};
    const int sv[6] = {SF_BRDF, SF_LENS, SF_LIGHT_UV, SF_LIGHT_SELECT, SF_ENV, SF_RR};
    const int sc[6] = {SF_BRDF_CONF, SF_LENS_CONF, SF_LIGHT_UV_CONF, SF_LIGHT_SELECT_CONF, SF_ENV_CONF, SF_RR_CONF};
    tiles[SF_PIXEL * nn + idx] = clampf(fin[0], 0.0f, 1.0f);
    tiles[SF_PIXEL_CONF * nn + idx] = clampf(ns[GRID_ACC * nn + idx], 0.0f, 1.0f);
    float rel[6], rc[6];
    for (int k = 0; k < 6; ++k) {
        const float rv = clampf(fin[rnd[k]], 0.0f, 1.0f);
        rc[k] = clampf(ns[(GRID_ACC + rnd[k]) * nn + idx], 0.0f, 1.0f);
        const float relative = rv / fmaxf(RPD_EPS_SMALL, rv + fin[scr[k]]);
        rel[k] = relative * fminf(rc[k], ns[(GRID_ACC + scr[k]) * nn + idx]);
        tiles[sc[k] * nn + idx] = rc[k];
    }
    int dom = 0;
    for (int k = 1; k < 6; ++k) if (rel[dom] < rel[k]) dom = k;
    for (int k = 0; k < 6; ++k) tiles[sv[k] * nn + idx] = rel[k];
    tiles[SF_LIGHT * nn + idx] = fmaxf(rel[2], rel[3]);
    tiles[SF_LIGHT_CONF * nn + idx] = rel[2] >= rel[3] ? rc[2] : rc[3];
    tiles[SF_ALL * nn + idx] = rel[dom];
    tiles[SF_ALL_CONF * nn + idx] = fminf(rc[dom], ns[(GRID_ACC + scr[dom]) * nn + idx]);
}

// core/buffers.h:68-86: bilinear lookup of a node field at pixel (x, y)
__device__ __forceinline__ float tile_lookup(const float* t, GridGeom g, int x, int y) {
    const float fx = float(x) / float(g.step), fy = float(y) / float(g.step);
    const int x0 = min(max(int(floorf(fx)), 0), g.nx - 1), y0 = min(max(int(floorf(fy)), 0), g.ny - 1);
    const int x1 = min(x0 + 1, g.nx - 1), y1 = min(y0 + 1, g.ny - 1);
    const float tx = clampf(fx - float(x0), 0.0f, 1.0f), ty = clampf(fy - float(y0), 0.0f, 1.0f);
    const float top = lerpf_(t[y0 * g.nx + x0], t[y0 * g.nx + x1], tx);
    const float bot = lerpf_(t[y1 * g.nx + x0], t[y1 * g.nx + x1], tx);
    return lerpf_(top, bot, ty);
}

// main.cpp:930-947
__global__ void k_gradient(Batch b) {
    const int64_t item = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (item >= b.total[BK_PIXEL]) return;
    int64_t li;
    const ImageDesc& dsc = b.img[batch_find(b, BK_PIXEL, item, li)];
    const GridGeom g = dsc.g;
    const int W = dsc.W, H = dsc.H;
    const float* tiles = dsc.tiles;
    float* conf = dsc.conf;
    float* grad = dsc.grad;
    const int i = int(li);
    const int x = i % W, y = i / W;
    const float* all = tiles + SF_ALL * g.nx * g.ny;
    const float l = tile_lookup(all, g, max(0, x - 1), y), r = tile_lookup(all, g, min(W - 1, x + 1), y);
    const float u = tile_lookup(all, g, x, max(0, y - 1)), d = tile_lookup(all, g, x, min(H - 1, y + 1));
    const float gr = sqrtf((r - l) * (r - l) + (d - u) * (d - u));
    grad[i] = clampf(0.5f * gr, 0.0f, 1.0f);
    conf[i] = tile_lookup(tiles + SF_ALL_CONF * g.nx * g.ny, g, x, y);
}

__global__ void k_zero_f32(float* p, int64_t n) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.0f;
}

void launch_sensitivity(const Batch& b, hipStream_t st) {
    const int64_t nn = b.total[BK_NODE], np = b.total[BK_PIXEL];
    hipLaunchKernelGGL(k_node_split, dim3((nn + 127) / 128), dim3(128), 0, st, b);
    hipLaunchKernelGGL(k_prior, dim3((nn + 127) / 128), dim3(128), 0, st, b);
    hipLaunchKernelGGL(k_gradient, dim3((np + 255) / 256), dim3(256), 0, st, b);
}
void launch_sensitivity_zero(GridGeom g, int W, int H, float* tiles, float* conf, float* grad, hipStream_t st) {
    const int64_t nt = int64_t(SF_COUNT) * g.nx * g.ny, np = int64_t(W) * H;
    hipLaunchKernelGGL(k_zero_f32, dim3((nt + 255) / 256), dim3(256), 0, st, tiles, nt);
    hipLaunchKernelGGL(k_zero_f32, dim3((np + 255) / 256), dim3(256), 0, st, conf, np);
    hipLaunchKernelGGL(k_zero_f32, dim3((np + 255) / 256), dim3(256), 0, st, grad, np);
}

}  // namespace rpd
