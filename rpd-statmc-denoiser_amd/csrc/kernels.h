// Kernel parameter blocks and launch wrappers shared by the engine's kernel
// files and its host orchestration (engine.cpp).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "rpd_device.h"

namespace rpd {

// Per-sample record, structure of arrays indexed by sample slot. Slots of one
// pixel are contiguous and in sample order: slot = offsets[pixel] + s.
//   r0 = (radiance rgb, flag bits)          flags: REC_*
//   r1 = (primary normal xyz, hit distance)
//   r2 = (primary albedo rgb, 0)
//   r3 = (primary position xyz, 0)
//   r4 = (pixel_u xy, lens_u xy)           canonical [0,1) draws
//   r5 = (brdf_u xy, light_u xy)
//   r6 = (environment_u xy)
enum : uint32_t {
    REC_FINITE = 1u << 0,
    REC_HIT = 1u << 1,
    REC_PIXEL = 1u << 2,
    REC_LENS = 1u << 3,
    REC_LSEL = 1u << 4,
    REC_BRDF = 1u << 5,
    REC_LIGHT = 1u << 6,
    REC_ENV = 1u << 7,
    REC_RR = 1u << 8,
    REC_RR_SURVIVED = 1u << 9,
    REC_LIGHT_SHIFT = 16,  // bits 16..23: light index of the first light selection
};
struct Records {
    float4* r0; float4* r1; float4* r2; float4* r3; float4* r4; float4* r5; float2* r6;
    int64_t capacity;
};

struct Camera {
    float pos[3], right[3], up[3], llc[3], horiz[3], vert[3];
    float lens_radius;
};

struct PassPlan {            // which pixels a pass samples and where their slots are
    const int* counts;       // samples per pixel
    const int* offsets;      // first slot per pixel (exclusive scan of counts)
    const int* active;       // pixel ids with counts > 0 (render passes: in 8x8 tile order)
    const int* totals;       // device: [0] samples in the pass, [1] active pixels (0 if over capacity)
    int n_pixels;
};

// Wavefront trace (flat scenes): every path slot's lane state between the
// shading and intersection kernels, as WF_WORDS planes of `slots` words
// (word k of slot s at state[k * slots + s]); the query lists of one round
// (slot * 4 + query kind, at most 3 per slot); per-round counters (for round
// i, ctrl[i * WF_LISTS + l] counts the queries of list l and
// ctrl[(WF_MAX_ITERS + i) * WF_LISTS + l] is set while a slot of list l still
// runs); and each 64-slot group's pixel hand-out cursor.
constexpr int WF_WORDS = 51;
constexpr int WF_MAX_ITERS = 512;
struct WavefrontState {
    uint32_t* state;
    int* queue;
    int* ctrl;
    int* cursor;
    int slots;
};
// Query lists: 64 sub-lists, a 64-slot group g appending to sub-list g % 64,
// so that no more than a 64th of the waves add to any one counter.
constexpr int WF_LISTS = 64;
constexpr int WF_MAX_SLOTS = 1 << 22;
__host__ __device__ __forceinline__ int wf_groups(int slots) { return (slots + 63) / 64; }
__host__ __device__ __forceinline__ int wf_list_cap(int slots) {
    return 3 * 64 * ((wf_groups(slots) + WF_LISTS - 1) / WF_LISTS);
}
__host__ __device__ __forceinline__ int64_t wf_queue_words(int slots) { return int64_t(WF_LISTS) * wf_list_cap(slots); }
constexpr int WF_CTRL_WORDS = 2 * WF_MAX_ITERS * WF_LISTS;

struct TraceParams {
    DevScene scene;
    Camera cam;
    int width, height;
    int capture;             // record random parameters (StatMC with RPD on)
    PassPlan plan;
    const int* seed_n;       // per-pixel sample offset used for seeding (finite samples so far), or null for 0
    uint32_t base_seed;
    uint32_t* tape;          // ceil(n/64)*64 tapes of 624 words and a seed, layout of tape_base/tape_word
    int* work_counter;       // zero before launch
    unsigned long long* query_count;  // ray queries issued (closest + occlusion), accumulated; may be null
    Records rec;
    // extensions, kept last so that the fields above keep their offsets
    int max_depth;           // shading points per path, 0 = unbounded
    DevBvh bvh;              // nodes null: the flat loop over scene.tri_perm
    int tape_per_sample, tape_margin;  // tape words written per sample and beyond (tape_prefix_words)
    WavefrontState wf;       // path slots of the wavefront trace (flat scenes), or null state
};

// Per-pixel accumulators (structure of arrays, n_pixels each unless noted)
struct PixelBuffers {
    float* radiance;         // 3 planes: r, g, b sums over finite samples
    float* albedo;           // 3 planes: sums over primary hits
    float* normal;           // 3 planes
    float* world_pos;        // 3 planes
    float* depth;            // hit distance sum
    int* hit_count;
    int* sample_count;
    // Welford statistics (path_tracer/core/stats.h:6-41)
    float* lum_mean;
    float* lum_m2;
    int* n;
    float* color_mean;       // 3 planes
    float* sqrt_mean;        // 3 planes
    float* sqrt_m2;          // 3 planes
};

struct Features {            // path_tracer/main.cpp:94-122
    float* normal;           // 3 planes, unit where the sum was non-degenerate
    float* depth;
    float* albedo;           // 3 planes
};

// RPF grid: 440 doubles per node, accumulators in this order, each laid out as
// weight[bins], sum_y[bins], w_sum, w2_sum, y_sum, y2_sum:
//   pixel, brdf, screen_brdf, lens, screen_lens, light_uv, screen_light_uv,
//   environment, screen_environment, light_select(16 bins),
//   screen_light_select, rr(2 bins), screen_rr
// Per-pixel filter scratch: 4 floats of sqrt-domain variance of the mean, 7
// float4 planes of packed filter inputs, 2 double2 planes of Welch-Satterthwaite
// terms.
constexpr int FILTER_SCRATCH_FLOATS = 4 + 7 * 4 + 2 * 4;
constexpr int GRID_ACC = 13;
constexpr int GRID_NODE_DOUBLES = 440;
__host__ __device__ constexpr int grid_acc_bins(int k) { return k == 11 ? 2 : 16; }
__host__ __device__ constexpr int grid_acc_offset(int k) { return k <= 11 ? 36 * k : 36 * 11 + 8; }

struct GridGeom { int support, step, nx, ny; };

// Sensitivity tiles over grid nodes (path_tracer/core/buffers.h:18-66), field order:
enum : int {
    SF_ALL = 0, SF_ALL_CONF, SF_PIXEL, SF_PIXEL_CONF, SF_BRDF, SF_BRDF_CONF, SF_LENS, SF_LENS_CONF,
    SF_LIGHT, SF_LIGHT_CONF, SF_LIGHT_UV, SF_LIGHT_UV_CONF, SF_LIGHT_SELECT, SF_LIGHT_SELECT_CONF,
    SF_ENV, SF_ENV_CONF, SF_RR, SF_RR_CONF, SF_COUNT
};

struct FilterConfig {
    int radius;
    float normal_threshold, depth_threshold, compat_alpha, sigma_max, rpd_scale;
};
struct VarianceConfig {
    int radius, iterations;
    float normal_threshold, depth_threshold, compat_sigma, shrinkage_k, sigma_max;
};

// ---- edge-aware smoothing windows --------------------------------------------
// Variance smoothing (main.cpp:1166-1268) and importance smoothing
// (main.cpp:1331-1368) filter a per-pixel field over a small window, gated or
// weighted by the neighbours' normals, depths and primary hits. A block
// filters an SMOOTH_TW x SMOOTH_TH tile and first stages, for the tile and its
// halo of `radius` pixels, the unit normal and depth (one float4) and the
// field value with the hit count (one float2) in LDS, so a pixel is read from
// HBM about once per block instead of once per neighbour. Radii above
// SMOOTH_MAX_R read the neighbours from global memory instead.
constexpr int SMOOTH_TW = 32, SMOOTH_TH = 8;
constexpr int SMOOTH_MAX_R = 8;
inline size_t smooth_lds_bytes(int radius) {
    return radius > SMOOTH_MAX_R ? 0
                                 : size_t(SMOOTH_TW + 2 * radius) * size_t(SMOOTH_TH + 2 * radius) *
                                       (sizeof(float4) + sizeof(float2));
}
// A pixel's guide features and field value, from the staged tile or from
// global memory (ld == nullptr).
struct SmoothSrc {
    const float4* nd;   // staged (normal xyz, depth), or null
    const float2* vh;   // staged (value, hit count bits)
    int span, x0, y0;
    int W, n;
    const float* nrm;   // 3 planes
    const float* dep;
    const int* hits;
    const float* val;
    __device__ __forceinline__ void get(int x, int y, float4& g, float& v, int& hit) const {
        if (nd) {
            const int k = (y - y0) * span + (x - x0);
            g = nd[k];
            const float2 t = vh[k];
            v = t.x;
            hit = __float_as_int(t.y);
        } else {
            const int q = y * W + x;
            g = make_float4(nrm[q], nrm[n + q], nrm[2 * n + q], dep[q]);
            v = val[q];
            hit = hits[q];
        }
    }
};
// Called by every thread of the block (it holds a barrier); s: dynamic LDS of
// smooth_lds_bytes(R) bytes, or null for the global-memory form.
__device__ __forceinline__ SmoothSrc smooth_stage(int W, int H, int R, const PixelBuffers& pb, const Features& f,
                                                  const float* src, float4* s) {
    SmoothSrc o;
    o.W = W; o.n = W * H;
    o.nrm = f.normal; o.dep = f.depth; o.hits = pb.hit_count; o.val = src;
    o.span = SMOOTH_TW + 2 * R;
    o.x0 = int(blockIdx.x) * SMOOTH_TW - R;
    o.y0 = int(blockIdx.y) * SMOOTH_TH - R;
    o.nd = nullptr; o.vh = nullptr;
    if (!s) return o;
    const int sh = SMOOTH_TH + 2 * R, cnt = o.span * sh;
    float2* vh = reinterpret_cast<float2*>(s + cnt);
    for (int k = threadIdx.x; k < cnt; k += blockDim.x) {
        const int xx = o.x0 + k % o.span, yy = o.y0 + k / o.span;
        if (xx >= 0 && xx < W && yy >= 0 && yy < H) {
            const int q = yy * W + xx;
            s[k] = make_float4(f.normal[q], f.normal[o.n + q], f.normal[2 * o.n + q], f.depth[q]);
            vh[k] = make_float2(src[q], __int_as_float(pb.hit_count[q]));
        }
    }
    __syncthreads();
    o.nd = s;
    o.vh = vh;
    return o;
}

// ---- batches of images ------------------------------------------------------
// The denoising stages (per-pixel fold, features, grid gather, sensitivity,
// StatMC) run over a batch of images in one launch per stage; a single image
// is a batch of one. Work items of all images are laid out back to back: for
// each item kind an exclusive prefix over the images maps a global item to
// (image, local item).
struct ImageDesc {
    int W, H, N;
    GridGeom g;
    PassPlan plan;
    Records rec;
    PixelBuffers pb;
    Features feat;
    double* grid;            // GRID_NODE_DOUBLES per node
    float* node_scratch;     // 33 planes over nodes
    float* tiles;            // SF_COUNT planes over nodes
    float* conf;             // per-pixel sensitivity confidence
    float* grad;             // per-pixel sensitivity gradient
    float* filter_scratch;   // FILTER_SCRATCH_FLOATS per pixel: variance of the mean (3 planes + 1
                             // spare), 7 float4 planes, 2 double2 planes
    float *den, *unc, *alp, *vtot, *veff, *vrat;
    FilterConfig fc;
    float conf_n0;
    float* dbg;              // null, or 3 planes the window filter fills per pixel with a primary hit:
                             // accepted neighbours, weight sum, effective neighbours (main.cpp:991-1108)
    uint32_t acc_mask;       // bit k: grid accumulator k can receive samples (GRID_ACC_ALL: every one)
    double* pair_w;          // the StatMC pair pass's forward-window weights (statmc_forward_count(radius)
                             // planes of N doubles), or null where the image has no room for them
};
constexpr uint32_t GRID_ACC_ALL = (1u << GRID_ACC) - 1u;
enum : int { BK_PIXEL = 0, BK_NODE = 1, BK_TILE = 2, BK_KINDS = 3 };
constexpr int BATCH_QUEUES = 8;     // one work range per XCD for persistent kernels
// StatMC window tile shapes (width, height): STATMC_SHAPES[k], k < STATMC_SHAPE_COUNT;
// shape 0 (16 x 16) is the default
constexpr int STATMC_SHAPE_COUNT = 7;
extern const int STATMC_SHAPES[STATMC_SHAPE_COUNT][2];
int statmc_shape_of(int tile_w, int tile_h);   // index of a shape, -1 if it is not one
struct Batch {
    const ImageDesc* img;            // device array, n entries
    const int64_t* prefix;           // device: BK_KINDS rows of n + 1 exclusive prefix sums
    unsigned int* queues;            // device: BK_KINDS * BATCH_QUEUES work counters
    int n;
    int64_t total[BK_KINDS];         // host copies of the last prefix entries (launch sizing)
    int max_step;                    // host: largest grid step of the images (0: not known)
    int tile_w, tile_h;              // StatMC tile shape the BK_TILE row counts
    int max_radius;                  // host: largest StatMC window radius of the images
    int pairs;                       // host: StatMC as the symmetric pair and sum passes (a batch of one
                                     // image, which has pair_w)
    int width1, height1;             // host: that image's size (the sum pass's grid); 0 for a batch
                                     // whose BK_TILE row counts 16 x 16 tiles
};
// prefix rows of a batch of one: {0, items} per item kind
struct BatchPrefix1 { int64_t v[BK_KINDS * 2]; };
static_assert(BK_KINDS * BATCH_QUEUES <= 64 && BK_KINDS * 2 <= 64, "k_put_batch1 writes with one wave");
// writes a batch of one's device copies (descriptor, prefix rows, zeroed work counters) on the stream
void launch_put_batch1(ImageDesc* dd, const ImageDesc& d, int64_t* dp, const BatchPrefix1& pf, unsigned int* dq,
                       hipStream_t st);
// pairs in the forward half of a StatMC window of radius r (dy > 0, or dy == 0 and dx > 0)
__host__ __device__ constexpr int statmc_forward_count(int r) { return 2 * r * (r + 1); }
// radii the pair form takes: its table fits a render job's tapes and its sum
// pass's LDS tile holds the window
bool statmc_pairs_radius_ok(int r);
__host__ __device__ inline int64_t image_items(int kind, int W, int H, GridGeom g, int tile_w = 16, int tile_h = 16) {
    return kind == BK_PIXEL ? int64_t(W) * H
         : kind == BK_NODE  ? int64_t(g.nx) * g.ny
                            : int64_t((W + tile_w - 1) / tile_w) * ((H + tile_h - 1) / tile_h);
}
// batch of one: sets its BK_TILE prefix row to {0, tiles} on the stream
void launch_set_tile_prefix1(const Batch& b, int64_t tiles, hipStream_t st);
// image of global item `item` of `kind` and its local index
__device__ __forceinline__ int batch_find(const Batch& b, int kind, int64_t item, int64_t& local) {
    const int64_t* pf = b.prefix + kind * (b.n + 1);
    int lo = 0, hi = b.n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (pf[mid] <= item) lo = mid; else hi = mid - 1;
    }
    local = item - pf[lo];
    return lo;
}

// Persistent work distribution. The items of one kind are split into
// BATCH_QUEUES contiguous ranges; a block starts on range blockIdx.x % 8 (the
// XCD the dispatcher places it on, so neighbouring items share that XCD's L2)
// and moves to the following ranges when its own is used up. Called by one
// thread; `cursor` carries the range the block last took from.
__device__ inline int64_t batch_take(const Batch& b, int kind, int& cursor) {
    const int64_t total = b.total[kind];
    for (; cursor < BATCH_QUEUES; ++cursor) {
        const int q = (int(blockIdx.x) + cursor) % BATCH_QUEUES;
        const int64_t lo = total * q / BATCH_QUEUES, hi = total * (q + 1) / BATCH_QUEUES;
        if (lo >= hi) continue;
        const int64_t v = atomicAdd(b.queues + kind * BATCH_QUEUES + q, 1u);
        if (lo + v < hi) return lo + v;
    }
    return -1;
}

// Blocks of `kernel` (block_threads threads, no dynamic LDS) that fit on the
// current device at once: occupancy per CU times the CU count, at least 1.
int device_resident_blocks(const void* kernel, int block_threads);

// stable radix sort of (key, value) pairs on the low key_bits bits (sort_keys.hip)
size_t sort_pairs_u32_temp_bytes(int n);
hipError_t sort_pairs_u32(const uint32_t* keys_in, uint32_t* keys_out, const int* vals_in, int* vals_out, int n,
                          int key_bits, void* temp, size_t temp_bytes, hipStream_t st);

// ---- launch wrappers (each enqueues on `st`) ----
void launch_tape(const TraceParams& p, hipStream_t st);
void launch_trace(const TraceParams& p, int n_blocks, hipStream_t st);
// Wavefront trace of a flat scene: `iterations` rounds of shading (one bounce
// or camera ray per path slot, queries listed) and intersection (one lane per
// listed query), then the path-regeneration kernel resumes the slots still
// running and finishes the pass. Same records as launch_trace.
void launch_trace_wavefront(const TraceParams& p, int iterations, int finish_blocks, hipStream_t st);
int trace_block_size();
int trace_blocks_per_cu(const DevScene& s, const DevBvh& b, int max_depth);  // resident trace blocks per CU
// too large for the flat loop's LDS tables: traced through a BVH
bool trace_needs_bvh(int n_tris, int n_spheres, int n_materials, int n_lights);
// per-pixel fold of the current pass's records of every image (Welford statistics when statmc)
void launch_fold(const Batch& b, int statmc, hipStream_t st);
void launch_plan_uniform(int n_pixels, int count, int* counts, int* offsets, int* active, int* totals,
                         hipStream_t st);
// exclusive scan of counts -> offsets, compaction -> active; writes totals[0] = samples, totals[1] = active.
// A pass whose samples exceed `capacity` gets totals[1] = 0 and sets totals[2] = 1.
void launch_plan_scan(int n_pixels, const int* counts, int* offsets, int* active, int* totals,
                      int64_t* scratch, int64_t capacity, hipStream_t st);
void launch_fill_i32(int* p, int value, int64_t n, hipStream_t st);
int64_t plan_scan_scratch_elems(int n_pixels);
// rewrites the active list of a plan (counts set) in 8x8 tile order, the order
// the trace hands pixels out in and the fold walks them in; with long_from > 0
// the pixels of at least long_from samples come first (scratch:
// plan_tile_scratch_elems entries)
void launch_plan_tile_order(int width, int height, const int* counts, int* active, int64_t* scratch,
                            int long_from, hipStream_t st);
int64_t plan_tile_scratch_elems(int width, int height);

void launch_features(const Batch& b, hipStream_t st);
void launch_grid_gather(const Batch& b, hipStream_t st);
void launch_sensitivity(const Batch& b, hipStream_t st);
void launch_sensitivity_zero(GridGeom g, int width, int height, float* tiles, float* confidence, float* gradient,
                             hipStream_t st);
// StatMC in two steps: per-pixel packing of the filter inputs into the
// images' filter scratch, then the window kernel reading them.
void launch_statmc_prepare(const Batch& b, hipStream_t st);
void launch_statmc_window(const Batch& b, hipStream_t st);
// synthetic per-sample records (denoise-only workload), slot = pixel * spp + sample
struct SynthParams {
    uint32_t seed;
    int W, H, spp, region;
    float mu, sigma, ff_frac, ff_scale, rho;
};
void launch_synth(const SynthParams& sp, const Records& rec, hipStream_t st);
// zero the persistent-kernel work counters of a batch
void launch_batch_queues_reset(const Batch& b, hipStream_t st);
void launch_variance(int width, int height, const VarianceConfig& vc, const PixelBuffers& pb, const Features& f,
                     float* tmp, float* out, hipStream_t st);
struct AllocConfig {
    int adaptive_spp, base_samples, radius;
    float sigma_max;
    uint32_t seed;
    float* dbg_variance_source;   // null, or a plane for the clamped variance source (main.cpp:1303-1321)
};
// importance field (main.cpp:1295-1371) into `importance`
void launch_importance(int width, int height, const AllocConfig& ac, const PixelBuffers& pb, const Features& f,
                       const float* variance_denoised, float* importance, void* scratch, hipStream_t st);
// counts from the importance field (main.cpp:1373-1438); scratch holds alloc_scratch_bytes(n)
void launch_alloc_counts(int n_pixels, const AllocConfig& ac, const float* importance, int* counts, void* scratch,
                         hipStream_t st);
size_t alloc_scratch_bytes(int n_pixels);
void launch_finalize(int n, const PixelBuffers& pb, float* out_radiance, float* out_albedo, float* out_normal,
                     float* out_pos, float* out_depth, hipStream_t st);

// ---- image-quality metrics (shared/tools/eval_tools.cpp:120-285) ----------
// An RGB float image: channel c of pixel i at p[c * cs + i * ps]
// (interleaved: cs 1, ps 3; planar: cs N, ps 1).
struct MetricImage { const float* p; int64_t cs, ps; };
// pass A partial fields; the last two are maxima, the others sums
enum : int {
    MA_SUM_SQ = 0, MA_SUM_ABS = 1, MA_MEAN_REF = 2, MA_MEAN_TST = 5, MA_REF_ENERGY = 8, MA_GRAD_ERR = 9,
    MA_GRAD_REF = 10, MA_RELMSE = 11, MA_PEAK = 12, MA_LOG_PEAK = 13, MA_FIELDS = 14
};
// pass B partial fields (sums)
enum : int { MB_VAR_REF = 0, MB_VAR_TST = 3, MB_COV = 6, MB_SSIM = 9, MB_FIELDS = 10 };
struct MetricsPassB { double mean_ref[3], mean_tst[3], c1, c2; };
constexpr int METRICS_A_MAX_BLOCKS = 1024;   // pass A grid-strides over at most this many blocks
int64_t metrics_a_blocks(int W, int H);
int64_t metrics_b_blocks(int W, int H);
// each writes its reduced fields (MA_FIELDS or MB_FIELDS doubles) to `out`;
// `part` holds blocks x fields doubles
void launch_metrics_a(const MetricImage& ref, const MetricImage& tst, int W, int H, double* part, double* out,
                      hipStream_t st);
void launch_metrics_b(const MetricImage& ref, const MetricImage& tst, int W, int H, const MetricsPassB& pb,
                      double* part, double* out, hipStream_t st);

// ---- comparison denoisers (denoising/joint_bilateral.cpp, atrous_wavelet.cpp)
// Images are interleaved float RGB, row-major, as the reference's tools read them.
struct BilateralParams {
    const float* raw; const float* normal; const float* albedo;   // normal encoded as (n + 1) / 2
    float* gaussian; float* bilateral; float* joint;             // any may be null
    int width, height, radius;
    float sigma_c, sigma_n, sigma_p, sigma_a;
};
int bilateral_radius(float sigma_p);                              // max(1, ceil(3 sigma_p))
void launch_joint_bilateral(const BilateralParams& bp, hipStream_t st);
struct AtrousPass {
    const float* in; float* out;
    const float* unit;      // unit normals (3 per pixel)
    const float* albedo;
    float* lum;             // scratch: luminance of `in`, one per pixel
    int width, height, step;
    float sigma_c, sigma_n, sigma_a;
};
void launch_atrous_prepare(const float* normal, int64_t n, float* unit, hipStream_t st);
void launch_atrous_pass(const AtrousPass& ap, hipStream_t st);

}  // namespace rpd
