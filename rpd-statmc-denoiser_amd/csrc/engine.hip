// Host side of the engine: contexts, scenes, jobs and the pass pipeline of
// render() (path_tracer/main.cpp:1543-1564) driven on one HIP stream per job.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <functional>
#include <memory>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/rpd_engine.h"
#include "bvh_build.h"
#include "host_scene.h"
#include "kernels.h"

using namespace rpd;
using namespace rpd_host;

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
    g_err = msg;
    return code;
}

// Stream capture (the hipGraph of a job's adaptive passes) must not overlap
// with synchronous copies, allocator calls or stream/event creation in other
// threads: the runtime rejects those while any stream is capturing (a legacy
// stream operation would depend on the capturing stream). Captures only record
// a few launches, so they hold this lock for their duration, and every such
// call of the engine takes it too.
std::mutex g_capture_mu;

hipError_t sync_memcpy(void* dst, const void* src, size_t bytes, hipMemcpyKind kind) {
    std::lock_guard<std::mutex> lk(g_capture_mu);
    return hipMemcpy(dst, src, bytes, kind);
}
template <class T>
hipError_t dev_malloc(T** p, size_t bytes) {
    std::lock_guard<std::mutex> lk(g_capture_mu);
    return hipMalloc(reinterpret_cast<void**>(p), bytes);
}
hipError_t dev_free(void* p) {
    std::lock_guard<std::mutex> lk(g_capture_mu);
    return hipFree(p);
}
hipError_t stream_create(hipStream_t* s) {
    std::lock_guard<std::mutex> lk(g_capture_mu);
    return hipStreamCreateWithFlags(s, hipStreamNonBlocking);
}
hipError_t stream_destroy(hipStream_t s) {
    std::lock_guard<std::mutex> lk(g_capture_mu);
    return hipStreamDestroy(s);
}
hipError_t event_create(hipEvent_t* e) {
    std::lock_guard<std::mutex> lk(g_capture_mu);
    return hipEventCreate(e);
}
hipError_t event_destroy(hipEvent_t e) {
    std::lock_guard<std::mutex> lk(g_capture_mu);
    return hipEventDestroy(e);
}

#define HIP_TRY(expr)                                                                          \
    do {                                                                                       \
        hipError_t e_ = (expr);                                                                \
        if (e_ != hipSuccess) return fail(RPD_ERR_DEVICE, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

float lum(h3 c) { return 0.2126f * c.x + 0.7152f * c.y + 0.0722f * c.z; }

template <class T>
int dalloc(T** p, size_t count, std::vector<void*>& owned) {
    *p = nullptr;
    if (count == 0) count = 1;
    void* q = nullptr;
    const hipError_t e = dev_malloc(&q, count * sizeof(T));
    if (e != hipSuccess) return fail(RPD_ERR_OUT_OF_MEMORY, std::string("hipMalloc: ") + hipGetErrorString(e));
    owned.push_back(q);
    *p = static_cast<T*>(q);
    return RPD_OK;
}

// Device scratch of rpd_image_metrics_compute_noalloc: the metric kernels'
// per-block partial sums followed by the reduced fields. Every device the
// library runs on holds its own copy until the process ends. Host code never
// uses `scratch` itself, only the address hipGetSymbolAddress returns for it
// (rpd_context::scratch), taken once per context with its device current.
constexpr int NOALLOC_METRICS_MAX_TILES = 65536;      // 16x16 tiles: up to 4096 x 4096 pixels
constexpr size_t NOALLOC_METRICS_OUT = 32;           // reduced fields (MA_FIELDS, MB_FIELDS <= 32)
constexpr size_t NOALLOC_METRICS_PARTS =             // partials of the larger pass
    size_t(NOALLOC_METRICS_MAX_TILES) * MB_FIELDS > size_t(METRICS_A_MAX_BLOCKS) * MA_FIELDS
        ? size_t(NOALLOC_METRICS_MAX_TILES) * MB_FIELDS : size_t(METRICS_A_MAX_BLOCKS) * MA_FIELDS;
static __device__ __align__(256) double scratch[NOALLOC_METRICS_PARTS + NOALLOC_METRICS_OUT];

}  // namespace

struct rpd_context {
    int device = 0;
    int num_cus = 0;
    uint64_t budget = 0;
    std::atomic<uint64_t> in_use{0};
    std::mutex mu;
    // rpd_denoise_batch: descriptor arrays (grown on demand), stream and timing events
    std::mutex batch_mu;
    rpd::ImageDesc* b_img = nullptr;
    int64_t* b_prefix = nullptr;
    unsigned int* b_queues = nullptr;
    int b_cap = 0;
    // batched StatMC pair form: the pair tables of a group of images, and the
    // group's descriptor and prefix arrays (grown on demand)
    double* pw = nullptr;
    size_t pw_bytes = 0;
    rpd::ImageDesc* p_img = nullptr;
    int64_t* p_prefix = nullptr;
    int p_cap = 0;
    hipStream_t b_stream = nullptr;
    hipEvent_t b_ev[2] = {nullptr, nullptr};
    bool graph_passes = false;  // capture each job's adaptive-pass sequence as a hipGraph and replay it
    // image metrics: partials, reduced fields and (host-image calls) staged images, grown on demand
    std::mutex metrics_mu;
    double* m_part = nullptr;
    size_t m_part_cap = 0;
    double* m_out = nullptr;
    float* m_img = nullptr;
    size_t m_img_cap = 0;
    hipStream_t m_stream = nullptr;
    // device address of this device's `scratch` array (set by rpd_ctx_create)
    double* scratch = nullptr;
    // comparison denoisers: staging and scratch floats, grown on demand
    std::mutex filters_mu;
    float* f_buf = nullptr;
    size_t f_cap = 0;
    hipStream_t f_stream = nullptr;
    // device slabs of released jobs, by size, for the next jobs: hipMalloc and
    // hipFree of a job's few GB cost milliseconds, and hipFree waits for all
    // work queued on the device, which stalls the caller while other jobs run.
    // Each entry is (slab, insertion number); at most cache_cap bytes are kept,
    // the oldest freed first.
    std::mutex cache_mu;
    std::multimap<size_t, std::pair<void*, uint64_t>> cache;
    uint64_t cache_bytes = 0;
    uint64_t cache_seq = 0;
    uint64_t cache_cap = uint64_t(32) << 30;
    uint64_t held_limit = 0;  // a larger kept slab only while jobs hold at most this (0: no limit)
    int bvh_builder = RPD_BVH_HOST_SAH;  // for the scenes created next
    // StatMC window tile shapes: tuned per (radius, image-size class), or forced
    std::mutex tune_mu;
    std::map<uint64_t, int> tuned;
    int forced_shape = -1;
    bool debug_outputs = false;  // jobs created next keep the StatMC and allocation debug planes
    // tape words written before a pass per sample and beyond (tape_prefix_words);
    // 48 per sample covers a cornell_box sample of 7 bounces (C1 per job: 49.7 ms
    // with whole tapes, 48.7 at 64/96, 48.4 at 48/64, 59.2 at 32/32, where the
    // trace extends too many tapes itself)
    int tape_per_sample = 48, tape_margin = 64;
    // flat-scene passes run as a wavefront trace of ceil(wf_rounds * samples
    // per slot) shading/intersection rounds before the path-regeneration
    // kernel finishes them; 0 keeps the path-regeneration kernel alone
    float wf_rounds = 0.0f;
    // jobs created next run a pass's grid gather beside the next pass's trace
    // (second plan and record set, second stream)
    bool gather_overlap = true;
    // jobs' StatMC windows run as the symmetric pair and sum passes where the
    // job has room for the pair weights (render jobs, radius up to 13)
    bool statmc_pairs = true;
    // streams of released jobs, for the next jobs: creating a stream costs
    // host time, and with more hardware queues than the default (the
    // GPU_MAX_HW_QUEUES environment variable) a new stream can also set up a
    // new queue, a few milliseconds
    std::mutex streams_mu;
    std::vector<hipStream_t> streams;
};

namespace {
constexpr size_t STREAM_POOL_MAX = 64;
hipError_t job_stream_take(rpd_context* c, hipStream_t* s) {
    {
        std::lock_guard<std::mutex> lk(c->streams_mu);
        if (!c->streams.empty()) {
            *s = c->streams.back();
            c->streams.pop_back();
            return hipSuccess;
        }
    }
    return stream_create(s);
}
// s is idle (its job synchronised it)
void job_stream_give(rpd_context* c, hipStream_t s) {
    {
        std::lock_guard<std::mutex> lk(c->streams_mu);
        if (c->streams.size() < STREAM_POOL_MAX) {
            c->streams.push_back(s);
            return;
        }
    }
    stream_destroy(s);
}
}  // namespace

struct rpd_scene {
    rpd_context* ctx = nullptr;
    std::vector<void*> owned;
    DevScene dev{};
    DevBvh bvh{};          // device BVH, or null nodes for the flat loop
    int n_tris = 0, n_spheres = 0, n_materials = 0, n_lights = 0;
    // resident trace blocks per CU without / with a depth limit, set on first
    // use; jobs of several threads may share the scene
    std::atomic<int> trace_per_cu[2] = {{0}, {0}};
    int bvh_builder = RPD_BVH_HOST_SAH;  // which builder made the BVH (if the scene has one)
    // host copies for host-side queries (rpd_scene_query)
    std::vector<DevTriangle> h_tris;
    std::vector<DevSphere> h_spheres;
    std::vector<float4> h_perm;
    BvhHost h_bvh;
};

struct rpd_job {
    rpd_context* ctx = nullptr;
    rpd_scene* scene = nullptr;
    rpd_render_config cfg{};
    rpd::Camera cam{};
    uint32_t seed = 0;
    hipStream_t stream = nullptr;
    void* slab = nullptr;   // every device buffer of the job, carved from one allocation
    uint64_t bytes = 0;     // the slab's size
    int W = 0, H = 0, N = 0;
    GridGeom geom{};
    // device state
    PixelBuffers pb{};
    Features feat{};
    float *den = nullptr, *unc = nullptr, *alp = nullptr, *vtot = nullptr, *veff = nullptr, *vrat = nullptr;
    float *vden = nullptr, *vtmp = nullptr, *imp = nullptr, *sconf = nullptr, *sgrad = nullptr;
    float *out_rad = nullptr, *out_raw = nullptr, *out_alb = nullptr, *out_nrm = nullptr, *out_pos = nullptr,
          *out_dep = nullptr;
    int *counts = nullptr, *offsets = nullptr, *active = nullptr, *totals = nullptr, *work = nullptr;
    unsigned long long* query_count = nullptr;  // [0] ray queries over all passes
    int64_t* scan_scratch = nullptr;
    uint32_t* tape = nullptr;
    Records rec{};
    WavefrontState wf{};     // wavefront trace slots (flat scenes), or slots 0
    double* grid = nullptr;
    float *tiles = nullptr, *node_scratch = nullptr, *filter_scratch = nullptr;
    void* alloc_scratch = nullptr;
    float* dbg = nullptr;    // debug planes (neighbours, weight sum, n_eff, variance source), or null
    ImageDesc* b1_img = nullptr;        // device copies of the batch of one
    int64_t* b1_prefix = nullptr;
    unsigned int* b1_queues = nullptr;
    bool denoiser = false;   // created for denoising records: no scene, camera or tape
    bool wf_on = false;      // the slab holds wavefront trace slots (the context had it on at creation)
    // Gather overlap: the slab holds a second plan and record set, and the
    // grid gather of a pass runs on `stream2` beside the next pass's trace.
    // counts/offsets/active/totals, rec and batch1 are those of set cur_set.
    bool ov_on = false;
    int cur_set = 0;
    int* plan_sets[2] = {nullptr, nullptr};  // counts, offsets, active, totals: 3N + 8 ints each
    Records rec_sets[2]{};
    Batch batch_sets[2]{};
    ImageDesc* b2_img = nullptr;             // device copies of set 1's batch of one
    int64_t* b2_prefix = nullptr;
    unsigned int* b2_queues = nullptr;
    hipStream_t stream2 = nullptr;
    hipEvent_t ev_go = nullptr;              // stream: the next pass's trace may start
    hipEvent_t ev_gathered[2] = {nullptr, nullptr};  // stream2: the gather of set k's records ended
    bool gathered_pending[2] = {false, false};       // ev_gathered[k] recorded and not yet waited on
    int gather_waiting = -1;                 // set whose records wait for their gather, or -1
    bool capturing = false;  // the stream is being captured into a graph: no per-stage timing events
    int max_spp = 0;         // record capacity per pixel of a denoiser job
    // this job as an image of a batch (device copies for the batch of one)
    ImageDesc desc{};
    Batch batch1{};
    bool features_valid = false;
    bool finalized = false;
    bool tail_pending = false;  // the sampling passes ran; the denoise tail has not
    bool last_pass_statmc = false;
    int last_pass_capture = 0;
    // timing
    struct Span { int stage; hipEvent_t a, b; };
    std::vector<Span> spans;
    double stage_ms[11] = {0};
    // async full-pipeline run
    std::thread worker;
    std::atomic<int> done{0};
    int status = RPD_OK;
    std::string error;
};

namespace {

// ---- stage timing -----------------------------------------------------------
struct StageTimer {
    rpd_job* j;
    int stage;
    hipStream_t on;          // the stream the stage's work is queued on
    hipEvent_t a = nullptr;
    StageTimer(rpd_job* job, int s, hipStream_t stream = nullptr) : j(job), stage(s), on(stream ? stream : job->stream) {
        if (!j->capturing && event_create(&a) == hipSuccess) hipEventRecord(a, on);
    }
    ~StageTimer() {
        hipEvent_t b = nullptr;
        if (a && !j->capturing && event_create(&b) == hipSuccess) {
            hipEventRecord(b, on);
            j->spans.push_back({stage, a, b});
        }
    }
};
void collect_spans(rpd_job* j) {
    for (auto& s : j->spans) {
        float ms = 0.0f;
        if (hipEventSynchronize(s.b) == hipSuccess && hipEventElapsedTime(&ms, s.a, s.b) == hipSuccess) {
            j->stage_ms[s.stage] += ms;
            if (s.stage == 7) j->stage_ms[8] += 1.0;
        }
        event_destroy(s.a);
        event_destroy(s.b);
    }
    j->spans.clear();
}

// ---- config validation: core/config.h:116-195 -------------------------------
int validate(const rpd_render_config& c) {
    auto bad = [](const char* m) { return fail(RPD_ERR_INVALID_ARG, m); };
    if (c.image_width <= 0 || c.image_height <= 0) return bad("image_width and image_height must be positive");
    if ((int64_t)c.image_width * c.image_height > (1 << 26)) return bad("image too large (more than 2^26 pixels)");
    if (c.samples_per_pixel <= 0) return bad("samples_per_pixel must be positive");
    if (c.rpf_tile_size <= 0) return bad("rpf_tile_size must be positive");
    if (c.color_window_radius <= 0) return bad("color_window_radius must be positive");
    if (!std::isfinite(c.color_normal_threshold) || c.color_normal_threshold <= 0.0f || c.color_normal_threshold > 1.0f)
        return bad("color_normal_threshold must be finite and in (0,1]");
    if (!std::isfinite(c.color_depth_threshold) || c.color_depth_threshold < 0.0f)
        return bad("color_depth_threshold must be finite and non-negative");
    if (!std::isfinite(c.color_compat_alpha) || c.color_compat_alpha <= 0.0f || c.color_compat_alpha >= 1.0f)
        return bad("color_compat_alpha must be finite and in (0,1)");
    if (!std::isfinite(c.color_sigma_max) || c.color_sigma_max <= 0.0f) return bad("color_sigma_max must be finite and positive");
    if (c.var_window_radius <= 0) return bad("var_window_radius must be positive");
    if (!std::isfinite(c.var_normal_threshold) || c.var_normal_threshold <= 0.0f || c.var_normal_threshold > 1.0f)
        return bad("var_normal_threshold must be finite and in (0,1]");
    if (!std::isfinite(c.var_depth_threshold) || c.var_depth_threshold < 0.0f)
        return bad("var_depth_threshold must be finite and non-negative");
    if (!std::isfinite(c.var_compat_sigma) || c.var_compat_sigma <= 0.0f) return bad("var_compat_sigma must be finite and positive");
    if (!std::isfinite(c.var_shrinkage_k) || c.var_shrinkage_k <= 0.0f) return bad("var_shrinkage_k must be finite and positive");
    if (c.adaptive_base_samples < 0) return bad("adaptive_base_samples must be non-negative");
    if (c.adaptive_spp < 0) return bad("adaptive_spp must be non-negative");
    if (c.adaptive_base_samples > c.adaptive_spp) return bad("adaptive_base_samples cannot exceed adaptive_spp");
    if (!std::isfinite(c.adaptive_sigma_max) || c.adaptive_sigma_max <= 0.0f) return bad("adaptive_sigma_max must be finite and positive");
    if (c.adaptive_passes < 0) return bad("adaptive_passes must be non-negative");
    if (c.adaptive_importance_smoothing_radius < 0) return bad("adaptive_importance_smoothing_radius must be non-negative");
    if (c.var_iterations <= 0) return bad("var_iterations must be positive");
    if (!std::isfinite(c.rpf_shrinkage_scale) || c.rpf_shrinkage_scale < 0.0f) return bad("rpf_shrinkage_scale must be finite and non-negative");
    if (!std::isfinite(c.rpf_confidence_samples) || c.rpf_confidence_samples <= 0.0f) return bad("rpf_confidence_samples must be finite and positive");
    if (c.max_depth < 0) return bad("max_depth must be non-negative");
    return RPD_OK;
}

// ---- camera: core/camera.h:31-88 ---------------------------------------------
int make_camera(const rpd_camera& in, float aspect, rpd::Camera& out) {
    const h3 dir = get3(in.direction), up_in = get3(in.world_up);
    if (!std::isfinite(in.fov_degrees) || in.fov_degrees <= 0.0f || in.fov_degrees >= 180.0f)
        return fail(RPD_ERR_INVALID_ARG, "fov must be finite and in (0,180)");
    if (!(hsq(dir) > 1e-8f) || !(hsq(up_in) > 1e-8f))
        return fail(RPD_ERR_INVALID_ARG, "Camera vectors must be finite and direction/up must be nonzero");
    if (!std::isfinite(in.focus_distance) || in.focus_distance <= 0.0f)
        return fail(RPD_ERR_INVALID_ARG, "focus_distance must be finite and positive");
    if (!std::isfinite(in.aperture) || in.aperture < 0.0f)
        return fail(RPD_ERR_INVALID_ARG, "aperture must be finite and non-negative");
    const h3 fwd = hnorm(dir);
    h3 up_hint = hnorm(up_in);
    h3 right = hcross(fwd, up_hint);
    if (hsq(right) < 1e-8f) {
        h3 alt = (std::fabs(fwd.z) < 0.999f) ? H3(0, 0, 1) : H3(0, 1, 0);
        const h3 d = alt - up_hint;
        if (hsq(d) <= 1e-8f * 1e-8f * std::min(hsq(alt), hsq(up_hint))) alt = H3(1, 0, 0);
        up_hint = alt;
        right = hcross(fwd, up_hint);
    }
    const h3 r = hnorm(right);
    const h3 u = hnorm(hcross(r, fwd));
    const float theta = float(double(in.fov_degrees) * 3.14159265358979323846 / double(180.0f));
    const float hh = std::tan(0.5f * theta);
    const float hw = aspect * hh;
    const float fd = in.focus_distance;
    const h3 horiz = (2.0f * fd * hw) * r;
    const h3 vert = (2.0f * fd * hh) * u;
    const h3 pos = get3(in.position);
    const h3 llc = pos + fwd * fd - 0.5f * horiz - 0.5f * vert;
    put3(out.pos, pos); put3(out.right, r); put3(out.up, u);
    put3(out.llc, llc); put3(out.horiz, horiz); put3(out.vert, vert);
    out.lens_radius = 0.5f * in.aperture;
    return RPD_OK;
}

// ---- job pipeline -----------------------------------------------------------------
void set_error_from(rpd_job* j, int st) {
    if (st != RPD_OK && j->status == RPD_OK) { j->status = st; j->error = g_err; }
}

int capture_enabled(const rpd_job* j, int statmc) { return statmc && j->cfg.rpf_shrinkage_scale > 0.0f; }

TraceParams trace_params(rpd_job* j, int statmc) {
    TraceParams p{};
    p.scene = j->scene->dev;
    p.cam = j->cam;
    p.width = j->W; p.height = j->H;
    p.capture = capture_enabled(j, statmc);
    p.plan.counts = j->counts; p.plan.offsets = j->offsets; p.plan.active = j->active;
    p.plan.totals = j->totals; p.plan.n_pixels = j->N;
    p.seed_n = statmc ? j->pb.n : nullptr;
    p.base_seed = j->seed;
    p.tape = j->tape;
    p.max_depth = j->cfg.max_depth;
    p.bvh = j->scene->bvh;
    p.tape_per_sample = j->ctx->tape_per_sample;
    p.tape_margin = j->ctx->tape_margin;
    p.work_counter = j->work;
    p.query_count = j->query_count;
    p.rec = j->rec;
    p.wf = j->wf;
    return p;
}

// Wavefront rounds for a pass of this job, or 0 for the path-regeneration
// kernel alone: flat scenes with slots, rounds for the pass's mean samples
// per pixel (the base pass's count, or the adaptive budget per pixel).
int wavefront_rounds(const rpd_job* j, int stage) {
    if (j->ctx->wf_rounds <= 0.0f || j->scene->bvh.nodes || j->wf.slots <= 0) return 0;
    const int per_px = stage == 0 ? j->cfg.samples_per_pixel : std::max(1, j->cfg.adaptive_spp);
    const double r = std::ceil(double(j->ctx->wf_rounds) * double(per_px) * double(j->N) / double(j->wf.slots));
    return int(std::min<double>(r, WF_MAX_ITERS));
}

int trace_blocks(rpd_job* j) {
    const int limited = j->cfg.max_depth > 0 ? 1 : 0;
    int per_cu = j->scene->trace_per_cu[limited].load();
    if (per_cu <= 0) {
        per_cu = trace_blocks_per_cu(j->scene->dev, j->scene->bvh, j->cfg.max_depth);
        j->scene->trace_per_cu[limited].store(per_cu);
    }
    const int by_work = (j->N + trace_block_size() - 1) / trace_block_size();
    return std::max(1, std::min(by_work, per_cu * std::max(1, j->ctx->num_cus)));
}

// A render plan's active list in the order the trace hands pixels out in.
// long_from > 0: pixels of at least that many samples are handed out first.
void order_plan(rpd_job* j, int long_from = 0) {
    launch_plan_tile_order(j->W, j->H, j->counts, j->active, j->scan_scratch, long_from, j->stream);
}

// ---- gather overlap -----------------------------------------------------------
// A job whose passes gather into the RPF grid runs the gather of pass k on a
// second stream, started once pass k+1's tape kernel has run, so it shares the
// device with pass k+1's trace. The trace holds every compute unit while
// pixels are left to hand out (its waves take all but 32 of each SIMD's 512
// VGPRs); when the active list runs out, its waves end one by one while the
// lanes still holding a pixel finish their samples in sequence, and the
// gather's blocks take the units they leave. Passes alternate between two plan
// and record sets, so pass k+1 plans and traces into the set pass k's gather
// is not reading; gathers keep their pass order on the second stream, and the
// last pass gathers on the job's stream after all of them, so the grid's sums
// are the same as with every gather in line.

// Makes plan and record set s current (counts, offsets, active, totals, rec,
// batch1); the job's stream first waits for a gather still reading set s.
int use_set(rpd_job* j, int s) {
    if (j->gathered_pending[s]) {
        HIP_TRY(hipStreamWaitEvent(j->stream, j->ev_gathered[s], 0));
        j->gathered_pending[s] = false;
    }
    const size_t N = j->N;
    int* q = j->plan_sets[s];
    j->counts = q; j->offsets = q + N; j->active = q + 2 * N; j->totals = q + 3 * N;
    j->rec = j->rec_sets[s];
    j->batch1 = j->batch_sets[s];
    j->cur_set = s;
    return RPD_OK;
}

// The job's stream waits for every gather queued on the second stream.
int wait_gathers(rpd_job* j) {
    for (int s = 0; s < 2; ++s)
        if (j->gathered_pending[s]) {
            HIP_TRY(hipStreamWaitEvent(j->stream, j->ev_gathered[s], 0));
            j->gathered_pending[s] = false;
        }
    return RPD_OK;
}

int sync_event_create(hipEvent_t* e) {
    std::lock_guard<std::mutex> lk(g_capture_mu);
    HIP_TRY(hipEventCreateWithFlags(e, hipEventDisableTiming));
    return RPD_OK;
}

int ensure_overlap_stream(rpd_job* j) {
    if (!j->stream2) HIP_TRY(job_stream_take(j->ctx, &j->stream2));
    int st = RPD_OK;
    if (!j->ev_go && (st = sync_event_create(&j->ev_go))) return st;
    for (auto& e : j->ev_gathered)
        if (!e && (st = sync_event_create(&e))) return st;
    return RPD_OK;
}

// Launches the gather of the set that waits for it on the second stream,
// after the work queued so far on the job's stream (the next pass's plan and
// tape kernel).
int launch_waiting_gather(rpd_job* j) {
    const int k = j->gather_waiting;
    j->gather_waiting = -1;
    HIP_TRY(hipEventRecord(j->ev_go, j->stream));
    HIP_TRY(hipStreamWaitEvent(j->stream2, j->ev_go, 0));
    {
        StageTimer tg(j, 10, j->stream2);
        launch_batch_queues_reset(j->batch_sets[k], j->stream2);
        launch_grid_gather(j->batch_sets[k], j->stream2);
    }
    HIP_TRY(hipEventRecord(j->ev_gathered[k], j->stream2));
    j->gathered_pending[k] = true;
    HIP_TRY(hipGetLastError());
    return RPD_OK;
}

// Sampling pass with the plan already in counts/offsets/active/totals. With
// defer_gather the pass's grid gather waits for the next pass (gather overlap).
int pass_after_plan(rpd_job* j, int statmc, int stage, bool defer_gather = false) {
    StageTimer tm(j, stage);
    const TraceParams p = trace_params(j, statmc);
    launch_fill_i32(j->work, 0, 1, j->stream);
    launch_tape(p, j->stream);
    if (j->gather_waiting >= 0) {
        const int st = launch_waiting_gather(j);
        if (st) return st;
    }
    {
        StageTimer tk(j, 7);
        const int rounds = wavefront_rounds(j, stage);
        if (rounds > 0) launch_trace_wavefront(p, rounds, trace_blocks(j), j->stream);
        else launch_trace(p, trace_blocks(j), j->stream);
    }
    launch_fold(j->batch1, statmc, j->stream);
    if (p.capture && defer_gather) {
        j->gather_waiting = j->cur_set;
    } else if (p.capture) {
        const int st = wait_gathers(j);   // the grid's sums follow the pass order
        if (st) return st;
        StageTimer tg(j, 10);
        launch_batch_queues_reset(j->batch1, j->stream);
        launch_grid_gather(j->batch1, j->stream);
    }
    j->features_valid = false;
    j->last_pass_statmc = statmc != 0;
    j->last_pass_capture = p.capture;
    HIP_TRY(hipGetLastError());
    return RPD_OK;
}

void ensure_features(rpd_job* j) {
    if (j->features_valid) return;
    launch_features(j->batch1, j->stream);
    j->features_valid = true;
}

int do_recompute_sensitivity(rpd_job* j) {
    StageTimer tm(j, 2);
    if (j->cfg.rpf_shrinkage_scale <= 0.0f) {
        launch_sensitivity_zero(j->geom, j->W, j->H, j->tiles, j->sconf, j->sgrad, j->stream);
    } else {
        ensure_features(j);
        launch_sensitivity(j->batch1, j->stream);
    }
    HIP_TRY(hipGetLastError());
    return RPD_OK;
}

// The StatMC window kernel's tile shape for a batch: a shape forced on the
// context, one tuned earlier for this (largest radius, size class), or one
// tuned now. Tuning launches the window kernel once per shape on this batch,
// timed with events, and keeps the fastest; every run computes the whole
// output with the same bits, so the batch's result is complete whichever run
// came last. `set_shape` points the batch's BK_TILE row at a shape's tiles.
// Under stream capture (no events, no waiting) an untuned key takes shape 0.
uint64_t tune_key(const Batch& b) {
    int size_class = 0;
    for (int64_t px = b.total[BK_PIXEL]; px > 1; px >>= 1) ++size_class;
    return (uint64_t(b.pairs ? 1 : 0) << 16) | (uint64_t(b.max_radius) << 8) | uint64_t(size_class);
}

int run_statmc_window(rpd_context* ctx, Batch& b, hipStream_t st, bool can_time,
                      const std::function<int(Batch&, int)>& set_shape) {
    int shape = ctx->forced_shape;
    const uint64_t key = tune_key(b);
    if (shape < 0) {
        std::lock_guard<std::mutex> lk(ctx->tune_mu);
        auto it = ctx->tuned.find(key);
        if (it != ctx->tuned.end()) shape = it->second;
    }
    if (shape < 0 && can_time) {
        // every shape twice, first in order and then in reverse (the first
        // launch finds colder caches), each timed by the faster of its two runs
        constexpr int kRuns = 2 * STATMC_SHAPE_COUNT;
        auto shape_of_run = [](int r) { return r < STATMC_SHAPE_COUNT ? r : kRuns - 1 - r; };
        hipEvent_t ev[kRuns + 1];
        for (auto& e : ev) HIP_TRY(event_create(&e));
        int st2 = RPD_OK;
        HIP_TRY(hipEventRecord(ev[0], st));
        for (int r = 0; r < kRuns && !st2; ++r) {
            st2 = set_shape(b, shape_of_run(r));
            launch_batch_queues_reset(b, st);
            launch_statmc_window(b, st);
            if (!st2 && hipEventRecord(ev[r + 1], st) != hipSuccess) st2 = fail(RPD_ERR_DEVICE, "hipEventRecord failed");
        }
        int best = 0;
        float best_ms = 0.0f;
        if (!st2 && hipEventSynchronize(ev[kRuns]) == hipSuccess) {
            float shape_ms[STATMC_SHAPE_COUNT];
            bool timed[STATMC_SHAPE_COUNT] = {};
            for (int r = 0; r < kRuns; ++r) {
                float ms = 0.0f;
                if (hipEventElapsedTime(&ms, ev[r], ev[r + 1]) != hipSuccess) continue;
                const int k = shape_of_run(r);
                if (!timed[k] || ms < shape_ms[k]) shape_ms[k] = ms;
                timed[k] = true;
            }
            bool any = false;
            for (int k = 0; k < STATMC_SHAPE_COUNT; ++k)
                if (timed[k] && (!any || shape_ms[k] < best_ms)) { best = k; best_ms = shape_ms[k]; any = true; }
            std::lock_guard<std::mutex> lk(ctx->tune_mu);
            ctx->tuned.emplace(key, best);
        }
        for (auto& e : ev) event_destroy(e);
        HIP_TRY(hipGetLastError());
        return st2;
    }
    if (shape < 0) shape = 0;
    const int st2 = set_shape(b, shape);
    if (st2) return st2;
    launch_batch_queues_reset(b, st);
    launch_statmc_window(b, st);
    return RPD_OK;
}

int do_statmc(rpd_job* j) {
    StageTimer tm(j, 3);
    ensure_features(j);
    launch_statmc_prepare(j->batch1, j->stream);
    int st = RPD_OK;
    {
        StageTimer tw(j, 9);
        auto set_shape = [j](Batch& b, int k) -> int {
            b.tile_w = STATMC_SHAPES[k][0];
            b.tile_h = STATMC_SHAPES[k][1];
            b.total[BK_TILE] = image_items(BK_TILE, j->W, j->H, j->geom, b.tile_w, b.tile_h);
            launch_set_tile_prefix1(b, b.total[BK_TILE], j->stream);
            return RPD_OK;
        };
        // the pair and sum passes where the job has room for the pair weights
        Batch b = j->batch1;
        b.pairs = (j->ctx->statmc_pairs && j->desc.pair_w && b.n == 1) ? 1 : 0;
        b.width1 = j->W;
        b.height1 = j->H;
        st = run_statmc_window(j->ctx, b, j->stream, !j->capturing, set_shape);
    }
    if (st) return st;
    HIP_TRY(hipGetLastError());
    return RPD_OK;
}

void launch_job_variance(rpd_job* j, hipStream_t st) {
    VarianceConfig vc;
    vc.radius = j->cfg.var_window_radius;
    vc.iterations = j->cfg.var_iterations;
    vc.normal_threshold = j->cfg.var_normal_threshold;
    vc.depth_threshold = j->cfg.var_depth_threshold;
    vc.compat_sigma = j->cfg.var_compat_sigma;
    vc.shrinkage_k = j->cfg.var_shrinkage_k;
    vc.sigma_max = j->cfg.adaptive_sigma_max;
    launch_variance(j->W, j->H, vc, j->pb, j->feat, j->vtmp, j->vden, st);
}

int do_variance(rpd_job* j) {
    StageTimer tm(j, 4);
    ensure_features(j);
    launch_job_variance(j, j->stream);
    HIP_TRY(hipGetLastError());
    return RPD_OK;
}

int do_allocation(rpd_job* j) {
    StageTimer tm(j, 5);
    ensure_features(j);
    AllocConfig ac;
    ac.adaptive_spp = j->cfg.adaptive_spp;
    ac.base_samples = j->cfg.adaptive_base_samples;
    ac.radius = j->cfg.adaptive_importance_smoothing_radius;
    ac.sigma_max = j->cfg.adaptive_sigma_max;
    ac.seed = j->seed;
    ac.dbg_variance_source = j->dbg ? j->dbg + 3 * size_t(j->N) : nullptr;
    launch_importance(j->W, j->H, ac, j->pb, j->feat, j->vden, j->imp, j->alloc_scratch, j->stream);
    launch_alloc_counts(j->N, ac, j->imp, j->counts, j->alloc_scratch, j->stream);
    launch_plan_scan(j->N, j->counts, j->offsets, j->active, j->totals, j->scan_scratch, j->rec.capacity, j->stream);
    // an adaptive pass gives some pixels up to eight times the mean: those
    // with twice the mean or more go first
    order_plan(j, 2 * std::max(1, j->cfg.adaptive_spp));
    HIP_TRY(hipGetLastError());
    return RPD_OK;
}

AllocConfig alloc_config(const rpd_job* j) {
    AllocConfig ac;
    ac.adaptive_spp = j->cfg.adaptive_spp;
    ac.base_samples = j->cfg.adaptive_base_samples;
    ac.radius = j->cfg.adaptive_importance_smoothing_radius;
    ac.sigma_max = j->cfg.adaptive_sigma_max;
    ac.seed = j->seed;
    ac.dbg_variance_source = nullptr;
    return ac;
}

int finalize_on(rpd_job* j, bool statmc, hipStream_t st) {
    launch_finalize(j->N, j->pb, statmc ? j->out_raw : j->out_rad, j->out_alb, j->out_nrm, j->out_pos, j->out_dep, st);
    if (statmc) HIP_TRY(hipMemcpyAsync(j->out_rad, j->den, sizeof(float) * 3 * j->N, hipMemcpyDeviceToDevice, st));
    j->finalized = true;
    HIP_TRY(hipGetLastError());
    return RPD_OK;
}

int do_finalize(rpd_job* j, bool statmc) { return finalize_on(j, statmc, j->stream); }

int reset_state(rpd_job* j) {
    const size_t N = j->N;
    int st = wait_gathers(j);
    if (!st) st = use_set(j, 0);
    if (st) return st;
    j->gather_waiting = -1;
    // zero every accumulator (buffers.init, pixel_stats.assign, rpf_grid.reset)
    HIP_TRY(hipMemsetAsync(j->pb.radiance, 0, sizeof(float) * 13 * N, j->stream));
    HIP_TRY(hipMemsetAsync(j->pb.hit_count, 0, sizeof(int) * 2 * N, j->stream));
    HIP_TRY(hipMemsetAsync(j->pb.lum_mean, 0, sizeof(float) * 12 * N, j->stream));
    HIP_TRY(hipMemsetAsync(j->den, 0, sizeof(float) * 14 * N, j->stream));
    // the finalized outputs too: a run that does not write one (raw radiance
    // without StatMC) leaves it zero, as buffers.init does (core/buffers.h:121-140)
    HIP_TRY(hipMemsetAsync(j->out_rad, 0, sizeof(float) * 16 * N, j->stream));
    HIP_TRY(hipMemsetAsync(j->grid, 0, sizeof(double) * GRID_NODE_DOUBLES * j->geom.nx * j->geom.ny, j->stream));
    HIP_TRY(hipMemsetAsync(j->tiles, 0, sizeof(float) * SF_COUNT * j->geom.nx * j->geom.ny, j->stream));
    HIP_TRY(hipMemsetAsync(j->counts, 0, sizeof(int) * (3 * N + 8), j->stream));  // plan, totals, work counter
    HIP_TRY(hipMemsetAsync(j->query_count, 0, sizeof(unsigned long long) * 2, j->stream));
    if (j->dbg) HIP_TRY(hipMemsetAsync(j->dbg, 0, sizeof(float) * 4 * N, j->stream));
    j->features_valid = false;
    j->finalized = false;
    return RPD_OK;
}

// render(): main.cpp:1543-1564. The StatMC calls whose results the reference
// overwrites before use (statmc_denoise at 1548 and recompute_sensitivity at
// 736/1551 when adaptive passes run) are not repeated; every output is
// identical without them.
// The adaptive passes of main.cpp:1451-1538 (variance smoothing, allocation,
// sampling pass) are the same launches with the same arguments every pass:
// all per-pass sizes live in device memory. The first pass is captured into a
// hipGraph and every pass, the first included, is a replay of it.
int run_adaptive_passes_graph(rpd_job* j, int passes) {
    hipGraph_t g = nullptr;
    hipGraphExec_t ge = nullptr;
    int st = RPD_OK;
    hipError_t e = hipSuccess;
    {
        // captured work only launches kernels on the job's stream: nothing in
        // it takes the capture lock
        std::lock_guard<std::mutex> lk(g_capture_mu);
        j->capturing = true;
        if (hipStreamBeginCapture(j->stream, hipStreamCaptureModeThreadLocal) != hipSuccess) {
            j->capturing = false;
            return fail(RPD_ERR_DEVICE, "hipStreamBeginCapture failed");
        }
        st = do_variance(j);
        if (!st) st = do_allocation(j);
        if (!st) st = pass_after_plan(j, 1, 1);
        e = hipStreamEndCapture(j->stream, &g);
        j->capturing = false;
    }
    if (st || e != hipSuccess || !g) {
        if (g) hipGraphDestroy(g);
        return st ? st : fail(RPD_ERR_DEVICE, std::string("graph capture failed: ") + hipGetErrorString(e));
    }
    if (hipGraphInstantiate(&ge, g, nullptr, nullptr, 0) != hipSuccess) {
        hipGraphDestroy(g);
        return fail(RPD_ERR_DEVICE, "hipGraphInstantiate failed");
    }
    hipError_t le = hipSuccess;
    for (int i = 0; i < passes && le == hipSuccess; ++i) {
        StageTimer tm(j, 1);
        le = hipGraphLaunch(ge, j->stream);
    }
    const hipError_t se = hipStreamSynchronize(j->stream);
    hipGraphExecDestroy(ge);
    hipGraphDestroy(g);
    if (le != hipSuccess || se != hipSuccess)
        return fail(RPD_ERR_DEVICE, std::string("graph launch failed: ") + hipGetErrorString(le != hipSuccess ? le : se));
    return RPD_OK;
}

// The sampling part: base pass and adaptive passes (main.cpp:1543-1547). A job
// without StatMC is finished here; otherwise tail_pending is set and
// run_tail, or a batch of several jobs' tails, completes it.
int run_passes(rpd_job* j) {
    j->tail_pending = false;
    int st = reset_state(j);
    if (st) return st;
    const bool statmc = j->cfg.use_statmc != 0;
    if (!statmc) {
        launch_plan_uniform(j->N, j->cfg.samples_per_pixel, j->counts, j->offsets, j->active, j->totals, j->stream);
        order_plan(j);
        if ((st = pass_after_plan(j, 0, 0))) return st;
        return do_finalize(j, false);
    }
    launch_plan_uniform(j->N, j->cfg.samples_per_pixel, j->counts, j->offsets, j->active, j->totals, j->stream);
    order_plan(j);
    const bool passes = j->cfg.adaptive_passes > 0 && j->cfg.adaptive_spp > 0;
    const bool overlap = passes && j->ov_on && !j->ctx->graph_passes;
    if (overlap && (st = ensure_overlap_stream(j))) return st;
    if ((st = pass_after_plan(j, 1, 0, overlap))) return st;
    if (!passes && (st = do_recompute_sensitivity(j))) return st;
    if (passes && j->ctx->graph_passes) {
        if ((st = run_adaptive_passes_graph(j, j->cfg.adaptive_passes))) return st;
    } else {
        for (int i = 0; passes && i < j->cfg.adaptive_passes; ++i) {
            if ((st = do_variance(j))) return st;
            if (overlap && (st = use_set(j, 1 - j->cur_set))) return st;
            if ((st = do_allocation(j))) return st;
            if ((st = pass_after_plan(j, 1, 1, overlap && i + 1 < j->cfg.adaptive_passes))) return st;
        }
    }
    j->tail_pending = true;
    return RPD_OK;
}

// The denoise tail (main.cpp:1548-1564): sensitivity from the final grid,
// StatMC, variance denoise of the final statistics, outputs.
int run_tail(rpd_job* j) {
    int st = do_recompute_sensitivity(j);
    if (!st) st = do_statmc(j);
    if (!st) st = do_variance(j);
    if (!st) st = do_finalize(j, true);
    j->tail_pending = false;
    return st;
}

int run_pipeline(rpd_job* j) {
    const int st = run_passes(j);
    if (st || !j->tail_pending) return st;
    return run_tail(j);
}

// The job's image descriptor (everything a denoise-stage kernel reads) and
// the batch of one that the job's own stages launch with.
ImageDesc make_desc(const rpd_job* j) {
    ImageDesc d{};
    d.W = j->W; d.H = j->H; d.N = j->N;
    d.g = j->geom;
    d.plan.counts = j->counts; d.plan.offsets = j->offsets; d.plan.active = j->active;
    d.plan.totals = j->totals; d.plan.n_pixels = j->N;
    d.rec = j->rec;
    d.pb = j->pb;
    d.feat = j->feat;
    d.grid = j->grid;
    d.node_scratch = j->node_scratch;
    d.tiles = j->tiles;
    d.conf = j->sconf;
    d.grad = j->sgrad;
    d.filter_scratch = j->filter_scratch;
    d.den = j->den; d.unc = j->unc; d.alp = j->alp; d.vtot = j->vtot; d.veff = j->veff; d.vrat = j->vrat;
    d.fc.radius = j->cfg.color_window_radius;
    d.fc.normal_threshold = j->cfg.color_normal_threshold;
    d.fc.depth_threshold = j->cfg.color_depth_threshold;
    d.fc.compat_alpha = j->cfg.color_compat_alpha;
    d.fc.sigma_max = j->cfg.color_sigma_max;
    d.fc.rpd_scale = j->cfg.rpf_shrinkage_scale;
    d.conf_n0 = j->cfg.rpf_confidence_samples;
    d.dbg = j->dbg;
    // The StatMC pair weights take the tape buffer, which no stage after the
    // last pass reads (a later pass writes its tapes anew): a render job's
    // tapes hold MT_TAPE_WORDS words per pixel, the weights
    // statmc_forward_count(radius) doubles.
    d.pair_w = nullptr;
    const int r = j->cfg.color_window_radius;
    if (!j->denoiser && j->tape && statmc_pairs_radius_ok(r)) d.pair_w = reinterpret_cast<double*>(j->tape);
    // grid accumulators this job's samples can reach: a render job records no
    // environment sample without an environment and no lens sample without an
    // aperture (k_trace); records given by the caller may carry any
    d.acc_mask = GRID_ACC_ALL;
    if (j->scene && !j->denoiser) {
        const float* e = j->scene->dev.env;
        if (!(e[0] > 0.0f || e[1] > 0.0f || e[2] > 0.0f)) d.acc_mask &= ~((1u << 7) | (1u << 8));
        if (!(j->cam.lens_radius > 0.0f)) d.acc_mask &= ~((1u << 3) | (1u << 4));
    }
    return d;
}

int build_batch1(rpd_job* j, ImageDesc* dd, int64_t* dp, unsigned int* dq) {
    j->desc = make_desc(j);
    BatchPrefix1 pf;
    for (int k = 0; k < BK_KINDS; ++k) {
        pf.v[2 * k] = 0;
        pf.v[2 * k + 1] = image_items(k, j->W, j->H, j->geom);
        j->batch1.total[k] = pf.v[2 * k + 1];
    }
    j->batch1.max_step = j->geom.step;
    j->batch1.tile_w = STATMC_SHAPES[0][0];
    j->batch1.tile_h = STATMC_SHAPES[0][1];
    j->batch1.max_radius = j->cfg.color_window_radius;
    launch_put_batch1(dd, j->desc, dp, pf, dq, j->stream);
    HIP_TRY(hipGetLastError());
    j->batch1.img = dd;
    j->batch1.prefix = dp;
    j->batch1.queues = dq;
    j->batch1.n = 1;
    return RPD_OK;
}

// Job slabs are sized in classes of eight per doubling (at most 12.5% above
// the job's need, 1 MiB at least), so that slabs of released jobs fit later
// jobs of similar size.
size_t slab_class(size_t bytes) {
    const size_t min_slab = size_t(1) << 20;
    if (bytes <= min_slab) return min_slab;
    const int e = 63 - __builtin_clzll(uint64_t(bytes));
    const size_t step = size_t(1) << (e - 3);
    return (bytes + step - 1) / step * step;
}

// A cached slab of at least `want` bytes and at most half as much again (the
// smallest such), or null; *got receives its size. A larger slab than `want`
// is taken only while the bytes jobs hold stay within the context's limit.
void* cache_take(rpd_context* c, size_t want, size_t* got) {
    std::lock_guard<std::mutex> lk(c->cache_mu);
    size_t most = want + want / 2;
    if (c->held_limit) {
        const uint64_t held = c->in_use.load();
        most = held + want >= c->held_limit ? want : std::min<uint64_t>(most, c->held_limit - held);
    }
    auto it = c->cache.lower_bound(want);
    if (it == c->cache.end() || it->first > most) return nullptr;
    void* p = it->second.first;
    *got = it->first;
    c->cache_bytes -= it->first;
    c->cache.erase(it);
    return p;
}

// Frees cached slabs, oldest first, until at most `keep` bytes are cached.
void cache_trim_locked(rpd_context* c, uint64_t keep) {
    while (c->cache_bytes > keep && !c->cache.empty()) {
        auto old = c->cache.begin();
        for (auto it = c->cache.begin(); it != c->cache.end(); ++it)
            if (it->second.second < old->second.second) old = it;
        dev_free(old->second.first);
        c->cache_bytes -= old->first;
        c->cache.erase(old);
    }
}
void cache_trim(rpd_context* c, uint64_t keep) {
    std::lock_guard<std::mutex> lk(c->cache_mu);
    cache_trim_locked(c, keep);
}

// The layout of a job's buffers in its slab: each buffer at a 256-byte
// aligned offset. With base null it only adds up the size.
struct SlabLayout {
    char* base = nullptr;
    size_t off = 0;
    template <class T>
    void take(T** p, size_t count) {
        off = (off + 255) & ~size_t(255);
        if (base) *p = reinterpret_cast<T*>(base + off);
        off += std::max<size_t>(count, 1) * sizeof(T);
    }
};

// Places every device buffer of job j (W, H, geom, cfg, denoiser, max_spp
// and ctx->debug_outputs set) in L; fills j's pointers when L has a base.
void job_layout(rpd_job* j, SlabLayout& L) {
    const size_t N = j->N;
    const size_t nn = size_t(j->geom.nx) * j->geom.ny;
#define A(ptr, n) L.take(&(ptr), (n))
    float* sums = nullptr; A(sums, 13 * N);
    j->pb.radiance = sums; j->pb.albedo = sums + 3 * N; j->pb.normal = sums + 6 * N; j->pb.world_pos = sums + 9 * N;
    j->pb.depth = sums + 12 * N;
    int* ic = nullptr; A(ic, 2 * N);
    j->pb.hit_count = ic; j->pb.sample_count = ic + N;
    float* stats = nullptr; A(stats, 12 * N);
    j->pb.lum_mean = stats; j->pb.lum_m2 = stats + N; j->pb.n = reinterpret_cast<int*>(stats + 2 * N);
    j->pb.color_mean = stats + 3 * N; j->pb.sqrt_mean = stats + 6 * N; j->pb.sqrt_m2 = stats + 9 * N;
    float* fe = nullptr; A(fe, 7 * N);
    j->feat.normal = fe; j->feat.depth = fe + 3 * N; j->feat.albedo = fe + 4 * N;
    float* misc = nullptr; A(misc, 14 * N);
    j->den = misc; j->unc = misc + 3 * N; j->alp = misc + 4 * N; j->vtot = misc + 5 * N; j->veff = misc + 6 * N;
    j->vrat = misc + 7 * N; j->vden = misc + 8 * N; j->vtmp = misc + 9 * N; j->imp = misc + 10 * N;
    j->sconf = misc + 11 * N; j->sgrad = misc + 12 * N;
    float* outs = nullptr; A(outs, 16 * N);
    j->out_rad = outs; j->out_raw = outs + 3 * N; j->out_alb = outs + 6 * N; j->out_nrm = outs + 9 * N;
    j->out_pos = outs + 12 * N; j->out_dep = outs + 15 * N;
    int* plan = nullptr; A(plan, 3 * N + 8);
    j->counts = plan; j->offsets = plan + N; j->active = plan + 2 * N; j->totals = plan + 3 * N;
    j->work = plan + 3 * N + 4;
    A(j->scan_scratch, (size_t)std::max(plan_scan_scratch_elems(j->N), plan_tile_scratch_elems(j->W, j->H)));
    A(j->query_count, 2);
    // a denoiser job draws no random numbers: one placeholder word instead of the tapes
    const size_t tape_words = j->denoiser ? 1 : ((N + MT_GROUP - 1) / MT_GROUP) * MT_GROUP * MT_TAPE_WORDS;
    A(j->tape, tape_words);
    const int per_px = j->denoiser ? j->max_spp : std::max(j->cfg.samples_per_pixel, j->cfg.adaptive_spp + 1);
    const int64_t cap = int64_t(per_px) * int64_t(N);
    j->rec.capacity = cap;
    A(j->rec.r0, cap); A(j->rec.r1, cap); A(j->rec.r2, cap); A(j->rec.r3, cap);
    A(j->rec.r4, cap); A(j->rec.r5, cap); A(j->rec.r6, cap);
    j->plan_sets[0] = plan;
    j->rec_sets[0] = j->rec;
    if (j->ov_on) {
        // the second set holds adaptive passes only (at most adaptive_spp * N samples)
        int* plan2 = nullptr; A(plan2, 3 * N + 8);
        j->plan_sets[1] = plan2;
        Records& r = j->rec_sets[1];
        const int64_t cap2 = int64_t(j->cfg.adaptive_spp + 1) * int64_t(N);
        r.capacity = cap2;
        A(r.r0, cap2); A(r.r1, cap2); A(r.r2, cap2); A(r.r3, cap2);
        A(r.r4, cap2); A(r.r5, cap2); A(r.r6, cap2);
        A(j->b2_img, 1);
        A(j->b2_prefix, BK_KINDS * 2);
        A(j->b2_queues, BK_KINDS * BATCH_QUEUES);
    }
    A(j->grid, nn * GRID_NODE_DOUBLES);
    A(j->tiles, nn * SF_COUNT);
    A(j->node_scratch, nn * 33);
    A(j->filter_scratch, FILTER_SCRATCH_FLOATS * N);
    char* as = nullptr; A(as, alloc_scratch_bytes(j->N));
    j->alloc_scratch = as;
    if (j->ctx && j->ctx->debug_outputs) A(j->dbg, 4 * N);
    A(j->b1_img, 1);
    A(j->b1_prefix, BK_KINDS * 2);
    A(j->b1_queues, BK_KINDS * BATCH_QUEUES);
    // wavefront trace slots: render jobs created while the context has the
    // wavefront trace on, whose tapes an unsigned 32-bit word offset can
    // address (the trace stores that offset per slot)
    if (j->wf_on && !j->denoiser && tape_words < (size_t(1) << 32)) {
        const int slots = int(std::min<size_t>(N, size_t(WF_MAX_SLOTS)));
        j->wf.slots = slots;
        A(j->wf.state, size_t(WF_WORDS) * slots);
        A(j->wf.queue, size_t(wf_queue_words(slots)));
        A(j->wf.ctrl, size_t(WF_CTRL_WORDS));
        A(j->wf.cursor, size_t(2) * wf_groups(slots));
    }
#undef A
}

// One slab for all of the job's buffers: a cached one when one fits, else
// hipMalloc (when that fails the cache is emptied and it is tried once more).
// Contents are undefined either way (reset_state clears what must start at zero).
int job_alloc(rpd_job* j) {
    SlabLayout L;
    job_layout(j, L);
    const size_t want = slab_class(L.off);
    size_t got = 0;
    void* q = cache_take(j->ctx, want, &got);
    if (!q) {
        got = want;
        hipError_t e = dev_malloc(&q, want);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            cache_trim(j->ctx, 0);
            e = dev_malloc(&q, want);
        }
        if (e != hipSuccess) return fail(RPD_ERR_OUT_OF_MEMORY, std::string("hipMalloc: ") + hipGetErrorString(e));
    }
    j->slab = q;
    j->ctx->in_use += got;
    j->bytes = got;
    L.base = static_cast<char*>(q);
    L.off = 0;
    job_layout(j, L);
    const int st = build_batch1(j, j->b1_img, j->b1_prefix, j->b1_queues);
    if (st) return st;
    j->batch_sets[0] = j->batch1;
    j->cur_set = 0;
    if (j->ov_on) {
        // set 1's batch of one: the same image with the second plan and records
        ImageDesc d = j->desc;
        const size_t N = j->N;
        int* p2 = j->plan_sets[1];
        d.plan.counts = p2; d.plan.offsets = p2 + N; d.plan.active = p2 + 2 * N; d.plan.totals = p2 + 3 * N;
        d.rec = j->rec_sets[1];
        BatchPrefix1 pf;
        for (int k = 0; k < BK_KINDS; ++k) { pf.v[2 * k] = 0; pf.v[2 * k + 1] = j->batch1.total[k]; }
        launch_put_batch1(j->b2_img, d, j->b2_prefix, pf, j->b2_queues, j->stream);
        HIP_TRY(hipGetLastError());
        Batch b = j->batch1;
        b.img = j->b2_img; b.prefix = j->b2_prefix; b.queues = j->b2_queues;
        j->batch_sets[1] = b;
    }
    return RPD_OK;
}

void set_geometry(rpd_job* j) {
    j->geom.support = j->cfg.rpf_tile_size;
    j->geom.step = std::max(1, j->cfg.rpf_tile_size / 2);
    j->geom.nx = std::max(2, (j->W + j->geom.step - 1) / j->geom.step + 1);
    j->geom.ny = std::max(2, (j->H + j->geom.step - 1) / j->geom.step + 1);
}

// Jobs whose adaptive passes gather into the RPF grid (StatMC on, adaptive
// passes, records captured) can overlap those gathers with the next trace.
bool gather_overlap_applies(const rpd_render_config& c) {
    return c.use_statmc != 0 && c.adaptive_passes > 0 && c.adaptive_spp > 0 && c.rpf_shrinkage_scale > 0.0f;
}

// Device bytes of a job (its slab, without debug planes); a denoiser job
// (max_spp > 0) has no tape and room for max_spp records per pixel; wf: with
// wavefront trace slots.
size_t job_bytes(const rpd_render_config& c, int max_spp = 0, bool wf = false, bool ov = false) {
    rpd_job t;
    t.cfg = c;
    t.W = c.image_width; t.H = c.image_height; t.N = t.W * t.H;
    t.denoiser = max_spp > 0;
    t.max_spp = max_spp;
    t.wf_on = wf;
    t.ov_on = ov && max_spp <= 0;
    set_geometry(&t);
    SlabLayout L;
    job_layout(&t, L);
    return slab_class(L.off);
}

void job_free(rpd_job* j) {
    if (j->worker.joinable()) j->worker.join();
    if (j->stream2) hipStreamSynchronize(j->stream2);
    if (j->stream) hipStreamSynchronize(j->stream);
    for (auto& s : j->spans) { event_destroy(s.a); event_destroy(s.b); }
    j->spans.clear();
    // the stream is idle: the slab goes to the context's cache, which frees its
    // oldest slabs beyond the cap
    if (j->slab) {
        std::lock_guard<std::mutex> lk(j->ctx->cache_mu);
        j->ctx->cache.emplace(size_t(j->bytes), std::make_pair(j->slab, j->ctx->cache_seq++));
        j->ctx->cache_bytes += j->bytes;
        cache_trim_locked(j->ctx, j->ctx->cache_cap);
        j->ctx->in_use -= j->bytes;
    }
    j->slab = nullptr;
    j->bytes = 0;
    if (j->stream) job_stream_give(j->ctx, j->stream);
    j->stream = nullptr;
    if (j->stream2) job_stream_give(j->ctx, j->stream2);
    j->stream2 = nullptr;
    for (hipEvent_t* e : {&j->ev_go, &j->ev_gathered[0], &j->ev_gathered[1]}) {
        if (*e) event_destroy(*e);
        *e = nullptr;
    }
}

// reset: clear the accumulators now (a job run stage by stage); a job queued
// for its whole pipeline at once is cleared by run_passes, its first step.
int job_create(rpd_context* ctx, rpd_scene* scene, const rpd_render_config* cfg, const rpd_camera* cam, uint32_t seed,
               rpd_job** out, bool reset = true) {
    if (!ctx || !scene || !cfg || !cam || !out) return fail(RPD_ERR_INVALID_ARG, "null argument");
    *out = nullptr;
    if (scene->ctx != ctx) return fail(RPD_ERR_INVALID_ARG, "scene belongs to another context or to none (host-only)");
    int st = validate(*cfg);
    if (st) return st;
    HIP_TRY(hipSetDevice(ctx->device));
    auto j = std::make_unique<rpd_job>();
    j->ctx = ctx; j->scene = scene; j->cfg = *cfg; j->seed = seed;
    j->W = cfg->image_width; j->H = cfg->image_height; j->N = j->W * j->H;
    if ((st = make_camera(*cam, float(j->W) / float(j->H), j->cam))) return st;
    set_geometry(j.get());
    j->wf_on = ctx->wf_rounds > 0.0f && !scene->bvh.nodes;
    // BVH scenes keep their gathers in line: beside the BVH walk (168 VGPRs,
    // 48 KB of LDS stack per block) the gather slowed the trace by more than
    // it hid (C2 job 782 -> 789 ms)
    j->ov_on = ctx->gather_overlap && gather_overlap_applies(*cfg) && !scene->bvh.nodes;
    if (ctx->budget && ctx->in_use.load() + job_bytes(*cfg, 0, j->wf_on, j->ov_on) > ctx->budget)
        return fail(RPD_ERR_OUT_OF_MEMORY, "job does not fit the context's memory budget");
    HIP_TRY(job_stream_take(ctx, &j->stream));
    if ((st = job_alloc(j.get()))) { job_free(j.get()); return st; }
    if (reset && (st = reset_state(j.get()))) { job_free(j.get()); return st; }
    *out = j.release();
    return RPD_OK;
}

int job_create_denoiser(rpd_context* ctx, const rpd_render_config* cfg, int max_spp, rpd_job** out) {
    if (!ctx || !cfg || !out) return fail(RPD_ERR_INVALID_ARG, "null argument");
    *out = nullptr;
    if (max_spp <= 0) return fail(RPD_ERR_INVALID_ARG, "max_samples_per_pixel must be positive");
    int st = validate(*cfg);
    if (st) return st;
    HIP_TRY(hipSetDevice(ctx->device));
    auto j = std::make_unique<rpd_job>();
    j->ctx = ctx; j->scene = nullptr; j->cfg = *cfg; j->seed = 0;
    j->denoiser = true; j->max_spp = max_spp;
    j->W = cfg->image_width; j->H = cfg->image_height; j->N = j->W * j->H;
    set_geometry(j.get());
    if (ctx->budget && ctx->in_use.load() + job_bytes(*cfg, max_spp) > ctx->budget)
        return fail(RPD_ERR_OUT_OF_MEMORY, "job does not fit the context's memory budget");
    HIP_TRY(job_stream_take(ctx, &j->stream));
    if ((st = job_alloc(j.get()))) { job_free(j.get()); return st; }
    if ((st = reset_state(j.get()))) { job_free(j.get()); return st; }
    *out = j.release();
    return RPD_OK;
}

bool capture_records(const rpd_job* j) { return j->cfg.rpf_shrinkage_scale > 0.0f; }

// Fold of the loaded pass + RPF grid (main.cpp:707-717), then
// recompute_sensitivity and statmc_denoise, on the job's own stream.
int do_denoise_records(rpd_job* j) {
    {
        StageTimer tm(j, 0);
        launch_fold(j->batch1, 1, j->stream);
        if (capture_records(j)) {
            StageTimer tg(j, 10);
            launch_batch_queues_reset(j->batch1, j->stream);
            launch_grid_gather(j->batch1, j->stream);
        }
    }
    j->features_valid = false;
    j->last_pass_statmc = true;
    j->last_pass_capture = capture_records(j);
    int st = do_recompute_sensitivity(j);
    if (!st) st = do_statmc(j);
    HIP_TRY(hipGetLastError());
    return st;
}

// ---- host-side conversion helpers for reads/writes -----------------------------
int copy_planes_to_interleaved(rpd_job* j, const float* dsrc, int planes, float* dst) {
    std::vector<float> tmp(size_t(planes) * j->N);
    HIP_TRY(hipMemcpyAsync(tmp.data(), dsrc, tmp.size() * sizeof(float), hipMemcpyDeviceToHost, j->stream));
    HIP_TRY(hipStreamSynchronize(j->stream));
    for (int i = 0; i < j->N; ++i)
        for (int c = 0; c < planes; ++c) dst[size_t(i) * planes + c] = tmp[size_t(c) * j->N + i];
    return RPD_OK;
}
int copy_interleaved_to_planes(rpd_job* j, const float* src, int planes, float* ddst) {
    std::vector<float> tmp(size_t(planes) * j->N);
    for (int i = 0; i < j->N; ++i)
        for (int c = 0; c < planes; ++c) tmp[size_t(c) * j->N + i] = src[size_t(i) * planes + c];
    HIP_TRY(hipMemcpyAsync(ddst, tmp.data(), tmp.size() * sizeof(float), hipMemcpyHostToDevice, j->stream));
    HIP_TRY(hipStreamSynchronize(j->stream));
    return RPD_OK;
}
float host_lerp(float a, float b, float t) { return a + (b - a) * t; }
// SensitivityTiles::lookup (core/buffers.h:68-86) on the host for expanded reads
void expand_tiles(const rpd_job* j, const float* t, float* out) {
    const GridGeom& g = j->geom;
    for (int y = 0; y < j->H; ++y)
        for (int x = 0; x < j->W; ++x) {
            const float fx = float(x) / float(g.step), fy = float(y) / float(g.step);
            const int x0 = std::clamp(int(std::floor(fx)), 0, g.nx - 1), y0 = std::clamp(int(std::floor(fy)), 0, g.ny - 1);
            const int x1 = std::min(x0 + 1, g.nx - 1), y1 = std::min(y0 + 1, g.ny - 1);
            const float tx = std::clamp(fx - float(x0), 0.0f, 1.0f), ty = std::clamp(fy - float(y0), 0.0f, 1.0f);
            const float top = host_lerp(t[y0 * g.nx + x0], t[y0 * g.nx + x1], tx);
            const float bot = host_lerp(t[y1 * g.nx + x0], t[y1 * g.nx + x1], tx);
            out[size_t(y) * j->W + x] = host_lerp(top, bot, ty);
        }
}

}  // namespace

namespace {

hipError_t read_back_sync(void* dst, const void* src, size_t bytes) {
    return sync_memcpy(dst, src, bytes, hipMemcpyDeviceToHost);
}

// The scene's BVH built on the device from its uploaded primitives, with a
// host copy for host-side queries. A tree too deep for the traversal stack
// leaves pn null (its buffers freed) so that the caller can use the host build.
int device_bvh_build(rpd_scene* sc, const DevTriangle* pt, int n_tris, const DevSphere* ps, int n_spheres,
                     BvhHost& bvh, Bvh4Node*& pn, float4*& pslot, int*& pprim) {
    const int n = n_tris + n_spheres;
    int st;
    if ((st = dalloc(&pn, size_t(n), sc->owned)) || (st = dalloc(&pslot, size_t(n) * BVH_SLOT_FLOAT4, sc->owned)) ||
        (st = dalloc(&pprim, size_t(n), sc->owned)))
        return st;
    void* scratch = nullptr;
    hipStream_t s = nullptr;
    HIP_TRY(dev_malloc(&scratch, bvh4_device_scratch_bytes(n)));
    if (stream_create(&s) != hipSuccess) {
        dev_free(scratch);
        return fail(RPD_ERR_DEVICE, "stream for the BVH build");
    }
    int count = 0, depth = 0;
    const int r = build_bvh4_device(pt, n_tris, ps, n_spheres, pn, pslot, pprim, scratch, &count, &depth,
                                    &read_back_sync, s);
    stream_destroy(s);
    dev_free(scratch);
    if (r == -2) {
        for (void* p : {static_cast<void*>(pn), static_cast<void*>(pslot), static_cast<void*>(pprim)}) {
            sc->owned.erase(std::find(sc->owned.begin(), sc->owned.end(), p));
            dev_free(p);
        }
        pn = nullptr; pslot = nullptr; pprim = nullptr;
        return RPD_OK;
    }
    if (r == -1) return fail(RPD_ERR_INVALID_ARG, "scene has no primitives");
    if (r) return fail(RPD_ERR_DEVICE, std::string("device BVH build: ") + hipGetErrorString(hipError_t(r)));
    bvh.nodes.resize(size_t(count));
    bvh.slots.resize(size_t(n) * BVH_SLOT_FLOAT4);
    bvh.slot_prim.resize(size_t(n));
    bvh.depth = depth;
    HIP_TRY(read_back_sync(bvh.nodes.data(), pn, bvh.nodes.size() * sizeof(Bvh4Node)));
    HIP_TRY(read_back_sync(bvh.slots.data(), pslot, bvh.slots.size() * sizeof(float4)));
    HIP_TRY(read_back_sync(bvh.slot_prim.data(), pprim, bvh.slot_prim.size() * sizeof(int)));
    return RPD_OK;
}

}  // namespace

namespace rpd_internal {
// Sets the calling thread's rpd_last_error: the serving loop hands a failure
// caught on its scheduler thread to the thread that waits for the request.
int set_error(int code, const char* msg) { return fail(code, msg ? msg : ""); }
}  // namespace rpd_internal

extern "C" {

void rpd_render_config_defaults(rpd_render_config* c) {
    if (!c) return;
    std::memset(c, 0, sizeof(*c));
    c->samples_per_pixel = 1;
    c->rpf_tile_size = 8;
    c->color_window_radius = 7;
    c->color_normal_threshold = 0.5f; c->color_depth_threshold = 0.25f;
    c->color_compat_alpha = 0.05f; c->color_sigma_max = 6.0f;
    c->var_window_radius = 1;
    c->var_normal_threshold = 0.5f; c->var_depth_threshold = 0.25f;
    c->var_compat_sigma = 30.0f; c->var_shrinkage_k = 1e-3f;
    c->var_iterations = 2;
    c->adaptive_base_samples = 0; c->adaptive_spp = 0;
    c->adaptive_sigma_max = 3.0f;
    c->adaptive_passes = 1; c->adaptive_importance_smoothing_radius = 1;
    c->rpf_shrinkage_scale = 1.0f; c->rpf_confidence_samples = 128.0f;
}

const char* rpd_last_error(void) { return g_err.c_str(); }

int rpd_ctx_create(int device, rpd_context** out) {
    if (!out) return fail(RPD_ERR_INVALID_ARG, "null out");
    *out = nullptr;
    int count = 0;
    HIP_TRY(hipGetDeviceCount(&count));
    if (device < 0 || device >= count) return fail(RPD_ERR_INVALID_ARG, "device index out of range");
    HIP_TRY(hipSetDevice(device));
    auto c = std::make_unique<rpd_context>();
    c->device = device;
    hipDeviceProp_t prop;
    HIP_TRY(hipGetDeviceProperties(&prop, device));
    c->num_cus = prop.multiProcessorCount;
    // keep released slabs up to a quarter of the device, at most 64 GiB: the
    // largest jobs (2048^2 at 32 spp) take about 30 GB each, and a cap below two
    // of them frees and re-allocates slabs between jobs of mixed sizes
    c->cache_cap = std::min<uint64_t>(uint64_t(64) << 30, uint64_t(prop.totalGlobalMem) / 4);
    // set-up of the allocation-free metrics call: this device's scratch array
    // and the stream it runs on, so that the call itself creates nothing
    HIP_TRY(hipGetSymbolAddress(reinterpret_cast<void**>(&c->scratch), HIP_SYMBOL(scratch)));
    HIP_TRY(stream_create(&c->m_stream));
    *out = c.release();
    return RPD_OK;
}
int rpd_ctx_destroy(rpd_context* ctx) {
    if (!ctx) return RPD_OK;
    if (ctx->b_stream) hipStreamSynchronize(ctx->b_stream);
    if (ctx->b_img) dev_free(ctx->b_img);
    if (ctx->b_prefix) dev_free(ctx->b_prefix);
    if (ctx->pw) dev_free(ctx->pw);
    if (ctx->p_img) dev_free(ctx->p_img);
    if (ctx->p_prefix) dev_free(ctx->p_prefix);
    if (ctx->b_queues) dev_free(ctx->b_queues);
    for (hipEvent_t e : ctx->b_ev) if (e) event_destroy(e);
    if (ctx->b_stream) stream_destroy(ctx->b_stream);
    if (ctx->m_stream) hipStreamSynchronize(ctx->m_stream);
    if (ctx->m_part) dev_free(ctx->m_part);
    if (ctx->m_out) dev_free(ctx->m_out);
    if (ctx->m_img) dev_free(ctx->m_img);
    if (ctx->m_stream) stream_destroy(ctx->m_stream);
    if (ctx->f_stream) hipStreamSynchronize(ctx->f_stream);
    if (ctx->f_buf) dev_free(ctx->f_buf);
    if (ctx->f_stream) stream_destroy(ctx->f_stream);
    for (hipStream_t st : ctx->streams) stream_destroy(st);
    ctx->streams.clear();
    cache_trim(ctx, 0);
    delete ctx;
    return RPD_OK;
}
int rpd_ctx_set_memory_budget(rpd_context* ctx, uint64_t bytes) {
    if (!ctx) return fail(RPD_ERR_INVALID_ARG, "null context");
    ctx->budget = bytes;
    return RPD_OK;
}
int rpd_ctx_set_bvh_builder(rpd_context* ctx, int builder) {
    if (!ctx) return fail(RPD_ERR_INVALID_ARG, "null context");
    if (builder != RPD_BVH_HOST_SAH && builder != RPD_BVH_DEVICE_LBVH) return fail(RPD_ERR_INVALID_ARG, "unknown BVH builder");
    ctx->bvh_builder = builder;
    return RPD_OK;
}

int rpd_ctx_set_tape_prefix(rpd_context* ctx, int words_per_sample, int margin) {
    if (!ctx) return fail(RPD_ERR_INVALID_ARG, "null context");
    if (words_per_sample < 0 || margin < 0) return fail(RPD_ERR_INVALID_ARG, "tape prefix sizes must be non-negative");
    ctx->tape_per_sample = words_per_sample;
    ctx->tape_margin = margin;
    return RPD_OK;
}

int rpd_ctx_set_held_limit(rpd_context* ctx, uint64_t bytes) {
    if (!ctx) return fail(RPD_ERR_INVALID_ARG, "null context");
    std::lock_guard<std::mutex> lk(ctx->cache_mu);
    ctx->held_limit = bytes;
    return RPD_OK;
}

int rpd_ctx_set_buffer_cache(rpd_context* ctx, uint64_t bytes, uint64_t* cached) {
    if (!ctx) return fail(RPD_ERR_INVALID_ARG, "null context");
    {
        std::lock_guard<std::mutex> lk(ctx->cache_mu);
        ctx->cache_cap = bytes;
    }
    cache_trim(ctx, bytes);
    if (cached) {
        std::lock_guard<std::mutex> lk(ctx->cache_mu);
        *cached = ctx->cache_bytes;
    }
    return RPD_OK;
}

int rpd_scene_create(rpd_context* ctx, const rpd_triangle* tris, int n_tris, const rpd_sphere* spheres, int n_spheres,
                     const rpd_material* mats, int n_mats, const rpd_light* lights, int n_lights, const float env[3],
                     rpd_scene** out) {
    if (!out || n_tris < 0 || n_spheres < 0 || n_mats <= 0 || n_lights < 0 || !mats)
        return fail(RPD_ERR_INVALID_ARG, "invalid scene arguments");
    if ((n_tris && !tris) || (n_spheres && !spheres) || (n_lights && !lights))
        return fail(RPD_ERR_INVALID_ARG, "null primitive array");
    *out = nullptr;
    if (ctx) HIP_TRY(hipSetDevice(ctx->device));
    std::vector<DevTriangle> dt(n_tris);
    for (int i = 0; i < n_tris; ++i) {
        const rpd_triangle& t = tris[i];
        if (t.material_id < 0 || t.material_id >= n_mats) return fail(RPD_ERR_INVALID_ARG, "triangle material out of range");
        std::memcpy(dt[i].v0, t.v0, 12); std::memcpy(dt[i].v1, t.v1, 12); std::memcpy(dt[i].v2, t.v2, 12);
        std::memcpy(dt[i].n, t.normal, 12);
        dt[i].material = t.material_id;
        // core/triangle.h:23-25: an emitter has some |emission| above 1e-8
        const bool em = std::fabs(t.emission[0]) > 1e-8f || std::fabs(t.emission[1]) > 1e-8f || std::fabs(t.emission[2]) > 1e-8f;
        dt[i].emitter = em ? 1 : 0;
        dt[i].light = t.light_id;
        if (em && (t.light_id < 0 || t.light_id >= n_lights)) return fail(RPD_ERR_INVALID_ARG, "emissive triangle without a light");
        dt[i].pad = 0;
    }
    std::vector<DevSphere> ds(n_spheres);
    for (int i = 0; i < n_spheres; ++i) {
        const rpd_sphere& s = spheres[i];
        if (s.material_id < 0 || s.material_id >= n_mats) return fail(RPD_ERR_INVALID_ARG, "sphere material out of range");
        std::memcpy(ds[i].c, s.center, 12);
        ds[i].r = s.radius;
        ds[i].material = s.material_id;
        const bool em = std::fabs(s.emission[0]) > 1e-8f || std::fabs(s.emission[1]) > 1e-8f || std::fabs(s.emission[2]) > 1e-8f;
        ds[i].emitter = em ? 1 : 0;
        ds[i].light = s.light_id;
        if (em && (s.light_id < 0 || s.light_id >= n_lights)) return fail(RPD_ERR_INVALID_ARG, "emissive sphere without a light");
        ds[i].pad = 0;
    }
    std::vector<DevMaterial> dm(n_mats);
    for (int i = 0; i < n_mats; ++i) {
        std::memset(&dm[i], 0, sizeof(DevMaterial));
        dm[i].kind = mats[i].kind;
        dm[i].beckmann = mats[i].distribution == RPD_MICROFACET_BECKMANN;
        dm[i].roughness = mats[i].roughness;
        std::memcpy(dm[i].albedo, mats[i].albedo, 12);
        std::memcpy(dm[i].n1, mats[i].n1, 12);
        std::memcpy(dm[i].n2, mats[i].n2, 12);
    }
    std::vector<DevLight> dl(n_lights);
    std::vector<float> cdf(n_lights);
    for (int i = 0; i < n_lights; ++i) {
        const rpd_light& l = lights[i];
        DevLight& d = dl[i];
        std::memset(&d, 0, sizeof(DevLight));
        d.kind = l.kind;
        std::memcpy(d.p0, l.p0, 12); std::memcpy(d.p1, l.p1, 12); std::memcpy(d.p2, l.p2, 12);
        std::memcpy(d.emission, l.emission, 12);
        const h3 em = get3(l.emission);
        if (l.kind == RPD_LIGHT_RECT) {  // core/light.h:22-26
            const h3 c = hcross(get3(l.p1), get3(l.p2));
            d.area = std::sqrt(hsq(c));
            put3(d.normal, c / d.area);
            d.power = d.area * std::max(0.0f, lum(em));
        } else if (l.kind == RPD_LIGHT_AREA) {  // core/light.h:62-63
            d.area = 0.5f * std::sqrt(hsq(hcross(get3(l.p1) - get3(l.p0), get3(l.p2) - get3(l.p0))));
            std::memcpy(d.normal, l.normal, 12);
            d.power = d.area * std::max(0.0f, lum(em));
        } else if (l.kind == RPD_LIGHT_POINT || l.kind == RPD_LIGHT_DIRECTIONAL) {
            d.power = std::max(0.0f, lum(em));
        } else {
            return fail(RPD_ERR_INVALID_ARG, "unknown light kind");
        }
        cdf[i] = (i == 0) ? d.power : cdf[i - 1] + d.power;  // core/scene.h:31-37
    }
    auto sc = std::make_unique<rpd_scene>();
    sc->ctx = ctx;
    // triangle vertices in the three cyclic axis orders of the intersector
    std::vector<float4> perm(size_t(n_tris) * 9);
    for (int i = 0; i < n_tris; ++i) {
        for (int kz = 0; kz < 3; ++kz) {
            const int kx = kz == 2 ? 0 : kz + 1, ky = kx == 2 ? 0 : kx + 1;
            const float* v[3] = {dt[i].v0, dt[i].v1, dt[i].v2};
            float f[12] = {};
            for (int k = 0; k < 3; ++k) { f[3 * k] = v[k][kx]; f[3 * k + 1] = v[k][ky]; f[3 * k + 2] = v[k][kz]; }
            float4* e = &perm[tri_perm_entry(i, kz)];
            for (int q = 0; q < 3; ++q) e[q] = make_float4(f[4 * q], f[4 * q + 1], f[4 * q + 2], f[4 * q + 3]);
        }
    }
    // what shading reads of each triangle, one float4 each
    if (n_mats > TRI_ATTR_MAX_MATERIALS || n_lights > TRI_ATTR_MAX_LIGHTS)
        return fail(RPD_ERR_UNSUPPORTED, "more than 65536 materials or 32767 lights");
    std::vector<float4> attr(static_cast<size_t>(n_tris));
    for (int i = 0; i < n_tris; ++i) attr[i] = tri_attr_pack(dt[i]);
    // scenes too large for the flat LDS loop get a BVH over triangles and spheres
    BvhHost bvh;
    const bool use_bvh = trace_needs_bvh(n_tris, n_spheres, n_mats, n_lights);
    bool device_bvh = use_bvh && ctx && ctx->bvh_builder == RPD_BVH_DEVICE_LBVH;
    if (use_bvh && !device_bvh) {
        const int bst = build_bvh4(dt.data(), n_tris, ds.data(), n_spheres, bvh);
        if (bst == -2) return fail(RPD_ERR_UNSUPPORTED, "BVH deeper than the traversal stack allows");
        if (bst) return fail(RPD_ERR_INVALID_ARG, "scene has no primitives");
    }
    DevTriangle* pt = nullptr; DevSphere* ps = nullptr; DevMaterial* pm = nullptr; DevLight* pl = nullptr;
    float* pc = nullptr; float4* pp = nullptr; float4* pa = nullptr;
    Bvh4Node* pn = nullptr; float4* pslot = nullptr; int* pprim = nullptr;
    int st;
    if (ctx) {  // a null context keeps the scene on the host (rpd_scene_query only)
        if ((st = dalloc(&pt, dt.size(), sc->owned)) || (st = dalloc(&pp, perm.size(), sc->owned)) ||
            (st = dalloc(&pa, attr.size(), sc->owned)) ||
            (st = dalloc(&ps, ds.size(), sc->owned)) ||
            (st = dalloc(&pm, dm.size(), sc->owned)) || (st = dalloc(&pl, dl.size(), sc->owned)) ||
            (st = dalloc(&pc, cdf.size(), sc->owned))) {
            for (void* p : sc->owned) dev_free(p);
            return st;
        }
        if (n_tris) {
            HIP_TRY(sync_memcpy(pt, dt.data(), dt.size() * sizeof(DevTriangle), hipMemcpyHostToDevice));
            HIP_TRY(sync_memcpy(pp, perm.data(), perm.size() * sizeof(float4), hipMemcpyHostToDevice));
            HIP_TRY(sync_memcpy(pa, attr.data(), attr.size() * sizeof(float4), hipMemcpyHostToDevice));
        }
        if (n_spheres) HIP_TRY(sync_memcpy(ps, ds.data(), ds.size() * sizeof(DevSphere), hipMemcpyHostToDevice));
        if (device_bvh) {
            if ((st = device_bvh_build(sc.get(), pt, n_tris, ps, n_spheres, bvh, pn, pslot, pprim))) return st;
            if (!pn) {  // deeper than the traversal stack allows: the host's SAH tree instead
                device_bvh = false;
                const int bst = build_bvh4(dt.data(), n_tris, ds.data(), n_spheres, bvh);
                if (bst == -2) return fail(RPD_ERR_UNSUPPORTED, "BVH deeper than the traversal stack allows");
                if (bst) return fail(RPD_ERR_INVALID_ARG, "scene has no primitives");
            }
        }
        if (use_bvh && !device_bvh) {
            if ((st = dalloc(&pn, bvh.nodes.size(), sc->owned)) || (st = dalloc(&pslot, bvh.slots.size(), sc->owned)) ||
                (st = dalloc(&pprim, bvh.slot_prim.size(), sc->owned))) {
                for (void* p : sc->owned) dev_free(p);
                return st;
            }
            HIP_TRY(sync_memcpy(pn, bvh.nodes.data(), bvh.nodes.size() * sizeof(Bvh4Node), hipMemcpyHostToDevice));
            HIP_TRY(sync_memcpy(pslot, bvh.slots.data(), bvh.slots.size() * sizeof(float4), hipMemcpyHostToDevice));
            HIP_TRY(sync_memcpy(pprim, bvh.slot_prim.data(), bvh.slot_prim.size() * sizeof(int), hipMemcpyHostToDevice));
        }
        HIP_TRY(sync_memcpy(pm, dm.data(), dm.size() * sizeof(DevMaterial), hipMemcpyHostToDevice));
        if (n_lights) {
            HIP_TRY(sync_memcpy(pl, dl.data(), dl.size() * sizeof(DevLight), hipMemcpyHostToDevice));
            HIP_TRY(sync_memcpy(pc, cdf.data(), cdf.size() * sizeof(float), hipMemcpyHostToDevice));
        }
    }
    sc->dev.tris = pt; sc->dev.n_tris = n_tris; sc->dev.tri_perm = pp; sc->dev.tri_attr = pa;
    sc->bvh.nodes = pn; sc->bvh.slots = pslot; sc->bvh.slot_prim = pprim;
    sc->bvh_builder = device_bvh ? RPD_BVH_DEVICE_LBVH : RPD_BVH_HOST_SAH;
    sc->h_tris = dt; sc->h_spheres = ds; sc->h_perm = std::move(perm); sc->h_bvh = std::move(bvh);
    sc->dev.spheres = ps; sc->dev.n_spheres = n_spheres;
    sc->dev.materials = pm; sc->dev.n_materials = n_mats;
    sc->dev.lights = pl; sc->dev.n_lights = n_lights;
    sc->dev.light_cdf = pc;
    for (int k = 0; k < 3; ++k) sc->dev.env[k] = env ? env[k] : 0.0f;
    sc->dev.env_on = 0;
    sc->dev.lambert_only = n_mats > 0 && std::all_of(dm.begin(), dm.end(), [](const DevMaterial& m) {
        return m.kind == MAT_LAMBERT;
    }) ? 1 : 0;
    sc->n_tris = n_tris; sc->n_spheres = n_spheres; sc->n_materials = n_mats; sc->n_lights = n_lights;
    *out = sc.release();
    return RPD_OK;
}

int rpd_scene_builtin(rpd_context* ctx, const char* name, rpd_scene** out) {
    if (!name || !out) return fail(RPD_ERR_INVALID_ARG, "null argument");
    HostScene s;
    std::string err;
    if (int st = build_builtin_scene(name, s, err)) return fail(st, err);
    return rpd_scene_create(ctx, s.tris.data(), int(s.tris.size()), s.spheres.data(), int(s.spheres.size()),
                            s.mats.data(), int(s.mats.size()), s.lights.data(), int(s.lights.size()), s.env, out);
}

int rpd_builtin_scene_arrays(const char* name, rpd_triangle* tris, int* n_tris, rpd_sphere* spheres, int* n_spheres,
                             rpd_material* mats, int* n_mats, rpd_light* lights, int* n_lights, float env[3]) {
    if (!name || !n_tris || !n_spheres || !n_mats || !n_lights) return fail(RPD_ERR_INVALID_ARG, "null argument");
    HostScene s;
    std::string err;
    if (int st = build_builtin_scene(name, s, err)) return fail(st, err);
    const bool fits = (!tris || *n_tris >= int(s.tris.size())) && (!spheres || *n_spheres >= int(s.spheres.size())) &&
                      (!mats || *n_mats >= int(s.mats.size())) && (!lights || *n_lights >= int(s.lights.size()));
    if (!fits) return fail(RPD_ERR_INVALID_ARG, "an output array is smaller than the scene's");
    if (tris) std::copy(s.tris.begin(), s.tris.end(), tris);
    if (spheres) std::copy(s.spheres.begin(), s.spheres.end(), spheres);
    if (mats) std::copy(s.mats.begin(), s.mats.end(), mats);
    if (lights) std::copy(s.lights.begin(), s.lights.end(), lights);
    if (env) std::copy(s.env, s.env + 3, env);
    *n_tris = int(s.tris.size()); *n_spheres = int(s.spheres.size());
    *n_mats = int(s.mats.size()); *n_lights = int(s.lights.size());
    return RPD_OK;
}

int rpd_obj_load(const char* path, int material_id, rpd_triangle* tris, int* n_tris) {
    if (!path || !n_tris) return fail(RPD_ERR_INVALID_ARG, "null argument");
    std::vector<rpd_triangle> t;
    std::string err;
    if (int st = load_obj_triangles(path, material_id, t, err)) return fail(st, err);
    if (tris) {
        if (*n_tris < int(t.size())) return fail(RPD_ERR_INVALID_ARG, "output array smaller than the mesh");
        std::copy(t.begin(), t.end(), tris);
    }
    *n_tris = int(t.size());
    return RPD_OK;
}

int rpd_scene_query(const rpd_scene* s, int n, const float* origins, const float* dirs, const float* t_max,
                    const int* occlusion, int use_bvh, int* hit, float* t, int* index, int* is_sphere) {
    if (!s || n < 0 || (n > 0 && (!origins || !dirs || !t_max || !occlusion || !hit || !t || !index || !is_sphere)))
        return fail(RPD_ERR_INVALID_ARG, "null argument");
    if (use_bvh && s->h_bvh.nodes.empty()) return fail(RPD_ERR_STATE, "scene has no BVH (small scenes are looped over)");
    DevScene hs = s->dev;
    hs.tris = s->h_tris.data();
    hs.spheres = s->h_spheres.data();
    hs.tri_perm = s->h_perm.data();
    DevBvh hb;
    hb.nodes = s->h_bvh.nodes.empty() ? nullptr : s->h_bvh.nodes.data();
    hb.slots = s->h_bvh.slots.empty() ? nullptr : s->h_bvh.slots.data();
    hb.slot_prim = s->h_bvh.slot_prim.empty() ? nullptr : s->h_bvh.slot_prim.data();
    int stack[BVH_STACK];
    for (int i = 0; i < n; ++i) {
        const v3 o = mk3(origins[3 * i], origins[3 * i + 1], origins[3 * i + 2]);
        const v3 d = mk3(dirs[3 * i], dirs[3 * i + 1], dirs[3 * i + 2]);
        const HitRec h = use_bvh ? bvh_query(hb, o, d, t_max[i], occlusion[i] != 0, stack, 1)
                                 : trace_query(hs, hs.tri_perm, o, d, t_max[i], occlusion[i] != 0);
        hit[i] = h.hit; t[i] = h.t; index[i] = h.index; is_sphere[i] = h.sphere;
    }
    return RPD_OK;
}

int rpd_scene_bvh_builder(const rpd_scene* s, int* builder) {
    if (!s || !builder) return fail(RPD_ERR_INVALID_ARG, "null argument");
    *builder = s->bvh_builder;
    return RPD_OK;
}

int rpd_scene_bvh_info(const rpd_scene* s, int* nodes, int* depth) {
    if (!s) return fail(RPD_ERR_INVALID_ARG, "null scene");
    if (nodes) *nodes = int(s->h_bvh.nodes.size());
    if (depth) *depth = s->h_bvh.depth;
    return RPD_OK;
}

int rpd_scene_counts(const rpd_scene* s, int* nt, int* ns, int* nm, int* nl) {
    if (!s) return fail(RPD_ERR_INVALID_ARG, "null scene");
    if (nt) *nt = s->n_tris;
    if (ns) *ns = s->n_spheres;
    if (nm) *nm = s->n_materials;
    if (nl) *nl = s->n_lights;
    return RPD_OK;
}

int rpd_scene_destroy(rpd_scene* s) {
    if (!s) return RPD_OK;
    for (void* p : s->owned) dev_free(p);
    delete s;
    return RPD_OK;
}

int rpd_job_create(rpd_context* ctx, rpd_scene* scene, const rpd_render_config* cfg, const rpd_camera* cam,
                   uint32_t seed, rpd_job** out) {
    return job_create(ctx, scene, cfg, cam, seed, out);
}

int rpd_job_submit(rpd_context* ctx, rpd_scene* scene, const rpd_render_config* cfg, const rpd_camera* cam,
                   uint32_t seed, rpd_job** out) {
    int st = job_create(ctx, scene, cfg, cam, seed, out, false);
    if (st) return st;
    rpd_job* j = *out;
    const int dev = ctx->device;
    j->worker = std::thread([j, dev]() {
        hipSetDevice(dev);
        const int s = run_pipeline(j);
        set_error_from(j, s);
        hipStreamSynchronize(j->stream);
        j->done.store(1);
    });
    return RPD_OK;
}

int rpd_job_submit_passes(rpd_context* ctx, rpd_scene* scene, const rpd_render_config* cfg, const rpd_camera* cam,
                          uint32_t seed, rpd_job** out) {
    int st = job_create(ctx, scene, cfg, cam, seed, out, false);
    if (st) return st;
    rpd_job* j = *out;
    const int dev = ctx->device;
    j->worker = std::thread([j, dev]() {
        hipSetDevice(dev);
        const int s = run_passes(j);
        set_error_from(j, s);
        hipStreamSynchronize(j->stream);
        j->done.store(1);
    });
    return RPD_OK;
}

int rpd_job_poll(rpd_job* j, int* done) {
    if (!j || !done) return fail(RPD_ERR_INVALID_ARG, "null argument");
    *done = j->worker.joinable() ? j->done.load() : 1;
    return RPD_OK;
}

int rpd_job_wait(rpd_job* j) {
    if (!j) return fail(RPD_ERR_INVALID_ARG, "null job");
    if (j->worker.joinable()) j->worker.join();
    HIP_TRY(hipStreamSynchronize(j->stream));
    collect_spans(j);
    if (j->status != RPD_OK) return fail(j->status, j->error);
    return RPD_OK;
}

int rpd_job_release(rpd_job* j) {
    if (!j) return RPD_OK;
    job_free(j);
    delete j;
    return RPD_OK;
}

int rpd_job_sync(rpd_job* j) {
    if (!j) return fail(RPD_ERR_INVALID_ARG, "null job");
    if (j->worker.joinable()) j->worker.join();
    HIP_TRY(hipStreamSynchronize(j->stream));
    collect_spans(j);
    int over = 0;
    HIP_TRY(sync_memcpy(&over, j->totals + 2, sizeof(int), hipMemcpyDeviceToHost));
    if (over) return fail(RPD_ERR_STATE, "an adaptive pass asked for more samples than the record capacity");
    return RPD_OK;
}

int rpd_job_run_pass(rpd_job* j, const int* counts, int statmc) {
    if (!j) return fail(RPD_ERR_INVALID_ARG, "null job");
    if (!j->scene) return fail(RPD_ERR_STATE, "job has no scene (it was created for denoising records)");
    if (int st0 = use_set(j, 0)) return st0;
    if (!counts) {
        launch_plan_uniform(j->N, j->cfg.samples_per_pixel, j->counts, j->offsets, j->active, j->totals, j->stream);
    } else {
        for (int i = 0; i < j->N; ++i)
            if (counts[i] < 0) return fail(RPD_ERR_INVALID_ARG, "negative sample count");
        HIP_TRY(hipMemcpyAsync(j->counts, counts, sizeof(int) * j->N, hipMemcpyHostToDevice, j->stream));
        launch_plan_scan(j->N, j->counts, j->offsets, j->active, j->totals, j->scan_scratch, j->rec.capacity, j->stream);
    }
    order_plan(j);
    const int st = pass_after_plan(j, statmc ? 1 : 0, counts ? 1 : 0);
    if (st) return st;
    return rpd_job_sync(j);
}

int rpd_job_run_adaptive_pass(rpd_job* j) {
    if (!j) return fail(RPD_ERR_INVALID_ARG, "null job");
    if (!j->scene) return fail(RPD_ERR_STATE, "job has no scene (it was created for denoising records)");
    int st = use_set(j, 0);
    if (!st) st = do_allocation(j);
    if (!st) st = pass_after_plan(j, 1, 1);
    if (st) return st;
    return rpd_job_sync(j);
}

int rpd_job_recompute_sensitivity(rpd_job* j) {
    if (!j) return fail(RPD_ERR_INVALID_ARG, "null job");
    const int st = do_recompute_sensitivity(j);
    return st ? st : rpd_job_sync(j);
}
int rpd_job_statmc_denoise(rpd_job* j) {
    if (!j) return fail(RPD_ERR_INVALID_ARG, "null job");
    const int st = do_statmc(j);
    return st ? st : rpd_job_sync(j);
}
int rpd_job_variance_denoise(rpd_job* j) {
    if (!j) return fail(RPD_ERR_INVALID_ARG, "null job");
    const int st = do_variance(j);
    return st ? st : rpd_job_sync(j);
}
int rpd_job_adaptive_counts(rpd_job* j, int* out) {
    if (!j) return fail(RPD_ERR_INVALID_ARG, "null job");
    int st = do_allocation(j);
    if (st) return st;
    if ((st = rpd_job_sync(j))) return st;
    if (out) HIP_TRY(sync_memcpy(out, j->counts, sizeof(int) * j->N, hipMemcpyDeviceToHost));
    return RPD_OK;
}
int rpd_job_adaptive_counts_from_importance(rpd_job* j, const float* importance, int* out) {
    if (!j || !importance) return fail(RPD_ERR_INVALID_ARG, "null argument");
    HIP_TRY(hipMemcpyAsync(j->imp, importance, sizeof(float) * j->N, hipMemcpyHostToDevice, j->stream));
    {
        StageTimer tm(j, 5);
        launch_alloc_counts(j->N, alloc_config(j), j->imp, j->counts, j->alloc_scratch, j->stream);
    }
    HIP_TRY(hipGetLastError());
    const int st = rpd_job_sync(j);
    if (st) return st;
    if (out) HIP_TRY(sync_memcpy(out, j->counts, sizeof(int) * j->N, hipMemcpyDeviceToHost));
    return RPD_OK;
}

int rpd_job_finalize(rpd_job* j) {
    if (!j) return fail(RPD_ERR_INVALID_ARG, "null job");
    const int st = do_finalize(j, j->cfg.use_statmc != 0);
    return st ? st : rpd_job_sync(j);
}

void rpd_synth_config_defaults(rpd_synth_config* c) {
    if (!c) return;
    c->seed = 1;
    c->samples_per_pixel = 16;
    c->region_size = 32;
    c->log_mean = -1.0f;
    c->log_sigma = 1.5f;
    c->firefly_fraction = 0.02f;
    c->firefly_scale = 100.0f;
    c->correlation = 0.6f;
}

int rpd_job_create_denoiser(rpd_context* ctx, const rpd_render_config* cfg, int max_samples_per_pixel,
                            rpd_job** out) {
    return job_create_denoiser(ctx, cfg, max_samples_per_pixel, out);
}

static_assert(sizeof(rpd_sample_record) == 100, "rpd_sample_record is 25 four-byte words");

int rpd_job_load_records(rpd_job* j, const int* counts, const rpd_sample_record* recs, size_t n) {
    if (!j || !counts || (n > 0 && !recs)) return fail(RPD_ERR_INVALID_ARG, "null argument");
    uint64_t total = 0;
    for (int i = 0; i < j->N; ++i) {
        if (counts[i] < 0) return fail(RPD_ERR_INVALID_ARG, "negative sample count");
        total += uint64_t(counts[i]);
    }
    if (total != n) return fail(RPD_ERR_INVALID_ARG, "record count does not match the sum of counts");
    if (int64_t(n) > j->rec.capacity) return fail(RPD_ERR_INVALID_ARG, "more records than the job's capacity");
    // records are given in pixel-major order, which is the slot order of the plan
    std::vector<float4> r[6];
    for (auto& v : r) v.resize(n);
    std::vector<float2> r6(n);
    for (size_t k = 0; k < n; ++k) {
        const rpd_sample_record& a = recs[k];
        uint32_t fl = a.flags & 0x3feu;  // the caller's validity bits; FINITE comes from the radiance
        if (std::isfinite(a.radiance[0]) && std::isfinite(a.radiance[1]) && std::isfinite(a.radiance[2]))
            fl |= REC_FINITE;
        fl |= uint32_t(std::clamp(a.light_index, 0, 255)) << REC_LIGHT_SHIFT;
        float fw;
        std::memcpy(&fw, &fl, 4);
        r[0][k] = make_float4(a.radiance[0], a.radiance[1], a.radiance[2], fw);
        r[1][k] = make_float4(a.normal[0], a.normal[1], a.normal[2], a.depth);
        r[2][k] = make_float4(a.albedo[0], a.albedo[1], a.albedo[2], 0.0f);
        r[3][k] = make_float4(a.position[0], a.position[1], a.position[2], 0.0f);
        r[4][k] = make_float4(a.pixel_u[0], a.pixel_u[1], a.lens_u[0], a.lens_u[1]);
        r[5][k] = make_float4(a.brdf_u[0], a.brdf_u[1], a.light_u[0], a.light_u[1]);
        r6[k] = make_float2(a.env_u[0], a.env_u[1]);
    }
    HIP_TRY(hipMemcpyAsync(j->counts, counts, sizeof(int) * j->N, hipMemcpyHostToDevice, j->stream));
    launch_plan_scan(j->N, j->counts, j->offsets, j->active, j->totals, j->scan_scratch, j->rec.capacity, j->stream);
    float4* dst[6] = {j->rec.r0, j->rec.r1, j->rec.r2, j->rec.r3, j->rec.r4, j->rec.r5};
    if (n) {
        for (int k = 0; k < 6; ++k)
            HIP_TRY(hipMemcpyAsync(dst[k], r[k].data(), n * sizeof(float4), hipMemcpyHostToDevice, j->stream));
        HIP_TRY(hipMemcpyAsync(j->rec.r6, r6.data(), n * sizeof(float2), hipMemcpyHostToDevice, j->stream));
    }
    j->last_pass_capture = 1;  // records carry their own validity bits
    HIP_TRY(hipGetLastError());
    return rpd_job_sync(j);
}

int rpd_job_synthesize_records(rpd_job* j, const rpd_synth_config* sc) {
    if (!j || !sc) return fail(RPD_ERR_INVALID_ARG, "null argument");
    if (sc->samples_per_pixel <= 0) return fail(RPD_ERR_INVALID_ARG, "samples_per_pixel must be positive");
    if (int64_t(sc->samples_per_pixel) * j->N > j->rec.capacity)
        return fail(RPD_ERR_INVALID_ARG, "more samples than the job's record capacity");
    if (sc->region_size <= 0) return fail(RPD_ERR_INVALID_ARG, "region_size must be positive");
    if (!(sc->correlation >= -1.0f && sc->correlation <= 1.0f)) return fail(RPD_ERR_INVALID_ARG, "correlation must be in [-1, 1]");
    launch_plan_uniform(j->N, sc->samples_per_pixel, j->counts, j->offsets, j->active, j->totals, j->stream);
    SynthParams sp;
    sp.seed = sc->seed; sp.W = j->W; sp.H = j->H; sp.spp = sc->samples_per_pixel; sp.region = sc->region_size;
    sp.mu = sc->log_mean; sp.sigma = sc->log_sigma; sp.ff_frac = sc->firefly_fraction; sp.ff_scale = sc->firefly_scale;
    sp.rho = sc->correlation;
    launch_synth(sp, j->rec, j->stream);
    j->last_pass_capture = 1;
    HIP_TRY(hipGetLastError());
    return rpd_job_sync(j);
}

int rpd_job_denoise_records(rpd_job* j) {
    if (!j) return fail(RPD_ERR_INVALID_ARG, "null job");
    const int st = do_denoise_records(j);
    return st ? st : rpd_job_sync(j);
}

// Checks a list of jobs for a batch call.
static int check_batch_list(rpd_context* ctx, rpd_job* const* jobs, int n) {
    if (!ctx || !jobs || n <= 0) return fail(RPD_ERR_INVALID_ARG, "null or empty batch");
    for (int k = 0; k < n; ++k) {
        if (!jobs[k]) return fail(RPD_ERR_INVALID_ARG, "null job in batch");
        if (jobs[k]->ctx != ctx) return fail(RPD_ERR_INVALID_ARG, "job of another context in batch");
        for (int m = 0; m < k; ++m)
            if (jobs[m] == jobs[k]) return fail(RPD_ERR_INVALID_ARG, "job listed twice in batch");
    }
    return RPD_OK;
}

// StatMC of a batch in the pair form (the symmetric pair pass and the in-order
// sum pass, the same bits as one pass), in groups of consecutive images whose
// pair tables fit one workspace of at most STATMC_PAIRS_WORKSPACE bytes; both
// passes take 16 x 16 tiles. Returns RPD_ERR_UNSUPPORTED, with nothing
// launched, when an image's radius does not take the pair form or no
// workspace holds its table: the caller then runs one pass.
constexpr size_t STATMC_PAIRS_WORKSPACE = size_t(16) << 30;
int batch_statmc_pairs(rpd_context* ctx, const std::vector<ImageDesc>& descs, const Batch& base, hipStream_t s,
                       std::vector<std::vector<int64_t>>& keep, std::vector<std::vector<ImageDesc>>& keep_d) {
    const int nb = int(descs.size());
    std::vector<size_t> bytes(nb);
    size_t most = 0, total = 0;
    for (int k = 0; k < nb; ++k) {
        if (!statmc_pairs_radius_ok(descs[k].fc.radius)) return RPD_ERR_UNSUPPORTED;
        bytes[k] = (size_t(descs[k].N) * statmc_forward_count(descs[k].fc.radius) * sizeof(double) + 255) / 256 * 256;
        most = std::max(most, bytes[k]);
        total += bytes[k];
    }
    const size_t want = std::min(total, std::max(most, STATMC_PAIRS_WORKSPACE));
    if (want > STATMC_PAIRS_WORKSPACE) return RPD_ERR_UNSUPPORTED;
    if (ctx->pw_bytes < want) {
        HIP_TRY(hipStreamSynchronize(s));   // earlier groups may still read the old workspace
        if (ctx->pw) HIP_TRY(dev_free(ctx->pw));
        ctx->pw = nullptr; ctx->pw_bytes = 0;
        if (dev_malloc(&ctx->pw, want) != hipSuccess) {
            (void)hipGetLastError();
            ctx->pw = nullptr;
            return RPD_ERR_UNSUPPORTED;
        }
        ctx->pw_bytes = want;
    }
    if (nb > ctx->p_cap) {
        HIP_TRY(hipStreamSynchronize(s));
        if (ctx->p_img) HIP_TRY(dev_free(ctx->p_img));
        if (ctx->p_prefix) HIP_TRY(dev_free(ctx->p_prefix));
        ctx->p_img = nullptr; ctx->p_prefix = nullptr; ctx->p_cap = 0;
        HIP_TRY(dev_malloc(&ctx->p_img, sizeof(ImageDesc) * nb));
        HIP_TRY(dev_malloc(&ctx->p_prefix, sizeof(int64_t) * BK_KINDS * (nb + 1)));
        ctx->p_cap = nb;
    }
    for (int g0 = 0; g0 < nb;) {
        int g1 = g0;
        size_t used = 0;
        while (g1 < nb && used + bytes[g1] <= ctx->pw_bytes) used += bytes[g1++];
        const int ng = g1 - g0;
        keep_d.emplace_back(descs.begin() + g0, descs.begin() + g1);
        std::vector<ImageDesc>& gd = keep_d.back();
        keep.emplace_back(size_t(BK_KINDS) * (ng + 1), 0);
        std::vector<int64_t>& pf = keep.back();
        Batch bg = base;
        bg.tile_w = bg.tile_h = 16;
        bg.max_radius = 0;
        size_t off = 0;
        for (int k = 0; k < ng; ++k) {
            gd[k].pair_w = reinterpret_cast<double*>(reinterpret_cast<char*>(ctx->pw) + off);
            off += bytes[g0 + k];
            bg.max_radius = std::max(bg.max_radius, gd[k].fc.radius);
        }
        for (int kind = 0; kind < BK_KINDS; ++kind) {
            for (int k = 0; k < ng; ++k)
                pf[size_t(kind) * (ng + 1) + k + 1] =
                    pf[size_t(kind) * (ng + 1) + k] + image_items(kind, gd[k].W, gd[k].H, gd[k].g, 16, 16);
            bg.total[kind] = pf[size_t(kind) * (ng + 1) + ng];
        }
        HIP_TRY(hipMemcpyAsync(ctx->p_img, gd.data(), sizeof(ImageDesc) * ng, hipMemcpyHostToDevice, s));
        HIP_TRY(hipMemcpyAsync(ctx->p_prefix, pf.data(), sizeof(int64_t) * pf.size(), hipMemcpyHostToDevice, s));
        bg.img = ctx->p_img; bg.prefix = ctx->p_prefix; bg.n = ng;
        bg.pairs = 1;
        bg.width1 = bg.height1 = 0;
        launch_batch_queues_reset(bg, s);
        launch_statmc_window(bg, s);
        HIP_TRY(hipGetLastError());
        g0 = g1;
    }
    return RPD_OK;
}

// The denoise stages of `jobs` as one batch on the context's batch stream
// (ctx->batch_mu held): with `records`, the fold of each job's loaded pass
// and the grid gather first (main.cpp:707-717); then features, sensitivity
// (main.cpp:741-953) and StatMC (main.cpp:957-1162), each one launch over
// every image.
static int batch_denoise_stages(rpd_context* ctx, const std::vector<rpd_job*>& jobs, bool records) {
    if (!ctx->b_stream) {
        HIP_TRY(stream_create(&ctx->b_stream));
        HIP_TRY(event_create(&ctx->b_ev[0]));
        HIP_TRY(event_create(&ctx->b_ev[1]));
        HIP_TRY(dev_malloc(&ctx->b_queues, sizeof(unsigned int) * BK_KINDS * BATCH_QUEUES));
    }
    const int nb = int(jobs.size());
    if (nb == 0) return RPD_OK;
    if (nb > ctx->b_cap) {
        if (ctx->b_img) HIP_TRY(dev_free(ctx->b_img));
        if (ctx->b_prefix) HIP_TRY(dev_free(ctx->b_prefix));
        ctx->b_img = nullptr; ctx->b_prefix = nullptr; ctx->b_cap = 0;
        HIP_TRY(dev_malloc(&ctx->b_img, sizeof(ImageDesc) * nb));
        HIP_TRY(dev_malloc(&ctx->b_prefix, sizeof(int64_t) * BK_KINDS * (nb + 1)));
        ctx->b_cap = nb;
    }
    hipStream_t s = ctx->b_stream;
    std::vector<ImageDesc> descs(nb);
    std::vector<int64_t> pf(size_t(BK_KINDS) * (nb + 1), 0);
    Batch b{};
    b.tile_w = STATMC_SHAPES[0][0];
    b.tile_h = STATMC_SHAPES[0][1];
    for (int k = 0; k < nb; ++k) {
        descs[k] = jobs[k]->desc;
        b.max_step = std::max(b.max_step, descs[k].g.step);
        b.max_radius = std::max(b.max_radius, descs[k].fc.radius);
    }
    for (int kind = 0; kind < BK_KINDS; ++kind) {
        for (int k = 0; k < nb; ++k)
            pf[size_t(kind) * (nb + 1) + k + 1] =
                pf[size_t(kind) * (nb + 1) + k] + image_items(kind, descs[k].W, descs[k].H, descs[k].g);
        b.total[kind] = pf[size_t(kind) * (nb + 1) + nb];
    }
    HIP_TRY(hipMemcpyAsync(ctx->b_img, descs.data(), sizeof(ImageDesc) * nb, hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(ctx->b_prefix, pf.data(), sizeof(int64_t) * pf.size(), hipMemcpyHostToDevice, s));
    b.img = ctx->b_img; b.prefix = ctx->b_prefix; b.queues = ctx->b_queues; b.n = nb;
    if (records) {
        launch_fold(b, 1, s);
        launch_batch_queues_reset(b, s);
        launch_grid_gather(b, s);
    }
    launch_features(b, s);
    launch_sensitivity(b, s);
    launch_statmc_prepare(b, s);
    // the BK_TILE prefix row for a tile shape, copied before the launch that uses it;
    // one host row per launch, all kept until the stream has drained below
    std::vector<std::vector<int64_t>> tile_rows;
    auto set_shape = [&](Batch& bb, int k) -> int {
        tile_rows.emplace_back(size_t(nb) + 1);
        std::vector<int64_t>& tile_row = tile_rows.back();
        bb.tile_w = STATMC_SHAPES[k][0];
        bb.tile_h = STATMC_SHAPES[k][1];
        tile_row[0] = 0;
        for (int m = 0; m < nb; ++m)
            tile_row[m + 1] = tile_row[m] + image_items(BK_TILE, descs[m].W, descs[m].H, descs[m].g, bb.tile_w,
                                                        bb.tile_h);
        bb.total[BK_TILE] = tile_row[nb];
        HIP_TRY(hipMemcpyAsync(ctx->b_prefix + size_t(BK_TILE) * (nb + 1), tile_row.data(),
                               sizeof(int64_t) * (nb + 1), hipMemcpyHostToDevice, s));
        return RPD_OK;
    };
    // the pair form where every image takes it and the workspace holds each table
    std::vector<std::vector<ImageDesc>> group_descs;
    int st = ctx->statmc_pairs ? batch_statmc_pairs(ctx, descs, b, s, tile_rows, group_descs) : RPD_ERR_UNSUPPORTED;
    if (st == RPD_ERR_UNSUPPORTED) st = run_statmc_window(ctx, b, s, true, set_shape);
    if (st) return st;
    // the host-side descriptor and prefix rows are read by asynchronous copies
    HIP_TRY(hipStreamSynchronize(s));
    HIP_TRY(hipGetLastError());
    for (rpd_job* j : jobs) {
        j->features_valid = true;
        j->last_pass_statmc = true;
        j->last_pass_capture = true;
    }
    return RPD_OK;
}

int rpd_denoise_batch(rpd_context* ctx, rpd_job* const* jobs, int n, double* gpu_ms) {
    int st = check_batch_list(ctx, jobs, n);
    if (st) return st;
    std::lock_guard<std::mutex> lk(ctx->batch_mu);
    HIP_TRY(hipSetDevice(ctx->device));
    // inputs were produced on the jobs' own streams
    std::vector<rpd_job*> batched, alone;
    for (int k = 0; k < n; ++k) {
        if (jobs[k]->worker.joinable()) jobs[k]->worker.join();
        HIP_TRY(hipStreamSynchronize(jobs[k]->stream));
        (capture_records(jobs[k]) ? batched : alone).push_back(jobs[k]);
    }
    if (!ctx->b_stream && (st = batch_denoise_stages(ctx, {}, true))) return st;  // creates the stream
    HIP_TRY(hipEventRecord(ctx->b_ev[0], ctx->b_stream));
    if ((st = batch_denoise_stages(ctx, batched, true))) return st;
    HIP_TRY(hipEventRecord(ctx->b_ev[1], ctx->b_stream));
    HIP_TRY(hipEventSynchronize(ctx->b_ev[1]));
    float ms = 0.0f;
    HIP_TRY(hipEventElapsedTime(&ms, ctx->b_ev[0], ctx->b_ev[1]));
    // images without RPF capture (rpf_shrinkage_scale <= 0) take the per-image path
    for (rpd_job* j : alone) {
        if ((st = rpd_job_denoise_records(j))) return st;
    }
    if (gpu_ms) *gpu_ms = double(ms);
    return RPD_OK;
}

int rpd_render_finish_batch(rpd_context* ctx, rpd_job* const* jobs, int n, double* gpu_ms) {
    int st = check_batch_list(ctx, jobs, n);
    if (st) return st;
    std::lock_guard<std::mutex> lk(ctx->batch_mu);
    HIP_TRY(hipSetDevice(ctx->device));
    std::vector<rpd_job*> batched, alone;
    for (int k = 0; k < n; ++k) {
        rpd_job* j = jobs[k];
        if (j->worker.joinable()) j->worker.join();
        HIP_TRY(hipStreamSynchronize(j->stream));
        if (j->status != RPD_OK) return fail(j->status, j->error);
        if (!j->tail_pending)
            return fail(RPD_ERR_STATE, "job has no pending denoise tail (start it with rpd_job_submit_passes)");
        (capture_records(j) ? batched : alone).push_back(j);
    }
    if (!ctx->b_stream && (st = batch_denoise_stages(ctx, {}, false))) return st;  // creates the stream
    hipStream_t s = ctx->b_stream;
    HIP_TRY(hipEventRecord(ctx->b_ev[0], s));
    if ((st = batch_denoise_stages(ctx, batched, false))) return st;
    for (rpd_job* j : batched) {          // main.cpp:1552-1564: variance denoise, outputs
        launch_job_variance(j, s);
        if ((st = finalize_on(j, true, s))) return st;
        j->tail_pending = false;
    }
    HIP_TRY(hipEventRecord(ctx->b_ev[1], s));
    HIP_TRY(hipEventSynchronize(ctx->b_ev[1]));
    float ms = 0.0f;
    HIP_TRY(hipEventElapsedTime(&ms, ctx->b_ev[0], ctx->b_ev[1]));
    // images without RPF capture take the per-image tail on their own streams
    for (rpd_job* j : alone) {
        if ((st = run_tail(j))) return st;
        HIP_TRY(hipStreamSynchronize(j->stream));
    }
    if (gpu_ms) *gpu_ms = double(ms);
    return RPD_OK;
}

int rpd_statmc_tile_shapes(int* wh, int n, int* count) {
    if (count) *count = STATMC_SHAPE_COUNT;
    for (int k = 0; wh && k < n && k < STATMC_SHAPE_COUNT; ++k) {
        wh[2 * k] = STATMC_SHAPES[k][0];
        wh[2 * k + 1] = STATMC_SHAPES[k][1];
    }
    return RPD_OK;
}

int rpd_ctx_set_statmc_tile(rpd_context* ctx, int tile_w, int tile_h) {
    if (!ctx) return fail(RPD_ERR_INVALID_ARG, "null context");
    if (tile_w == 0 && tile_h == 0) { ctx->forced_shape = -1; return RPD_OK; }
    const int k = statmc_shape_of(tile_w, tile_h);
    if (k < 0) return fail(RPD_ERR_INVALID_ARG, "not a StatMC tile shape (see rpd_statmc_tile_shapes)");
    ctx->forced_shape = k;
    return RPD_OK;
}

int rpd_ctx_statmc_tile(rpd_context* ctx, int radius, int64_t pixels, int* tile_w, int* tile_h) {
    if (!ctx || !tile_w || !tile_h) return fail(RPD_ERR_INVALID_ARG, "null argument");
    Batch b{};
    b.max_radius = radius;
    b.total[BK_PIXEL] = pixels;
    int k = ctx->forced_shape;
    if (k < 0) {   // the pair form's shape if it was tuned, else the one-pass form's
        std::lock_guard<std::mutex> lk(ctx->tune_mu);
        b.pairs = 1;
        auto it = ctx->tuned.find(tune_key(b));
        if (it == ctx->tuned.end()) {
            b.pairs = 0;
            it = ctx->tuned.find(tune_key(b));
        }
        if (it != ctx->tuned.end()) k = it->second;
    }
    *tile_w = k < 0 ? 0 : STATMC_SHAPES[k][0];
    *tile_h = k < 0 ? 0 : STATMC_SHAPES[k][1];
    return RPD_OK;
}

int rpd_ctx_set_debug_outputs(rpd_context* ctx, int on) {
    if (!ctx) return fail(RPD_ERR_INVALID_ARG, "null context");
    ctx->debug_outputs = on != 0;
    return RPD_OK;
}

int rpd_ctx_set_trace_wavefront(rpd_context* ctx, float rounds_per_sample) {
    if (!ctx) return fail(RPD_ERR_INVALID_ARG, "null context");
    if (!(rounds_per_sample >= 0.0f) || rounds_per_sample > 64.0f)
        return fail(RPD_ERR_INVALID_ARG, "rounds_per_sample must lie in [0, 64]");
    ctx->wf_rounds = rounds_per_sample;
    return RPD_OK;
}

int rpd_ctx_set_statmc_pairs(rpd_context* ctx, int on) {
    if (!ctx) return fail(RPD_ERR_INVALID_ARG, "null context");
    ctx->statmc_pairs = on != 0;
    return RPD_OK;
}

int rpd_ctx_set_gather_overlap(rpd_context* ctx, int on) {
    if (!ctx) return fail(RPD_ERR_INVALID_ARG, "null context");
    ctx->gather_overlap = on != 0;
    return RPD_OK;
}

int rpd_ctx_set_graph_passes(rpd_context* ctx, int on) {
    if (!ctx) return fail(RPD_ERR_INVALID_ARG, "null context");
    ctx->graph_passes = on != 0;
    return RPD_OK;
}

int rpd_abi_struct_sizes(int* out, int n) {
    if (!out) return fail(RPD_ERR_INVALID_ARG, "null argument");
    const int s[14] = {int(sizeof(rpd_render_config)), int(sizeof(rpd_camera)), int(sizeof(rpd_triangle)),
                       int(sizeof(rpd_sphere)), int(sizeof(rpd_material)), int(sizeof(rpd_light)),
                       int(sizeof(rpd_sample_record)), int(sizeof(rpd_synth_config)),
                       int(sizeof(rpd_procedural_config)), int(sizeof(rpd_server_config)),
                       int(sizeof(rpd_server_stats)), int(sizeof(rpd_image_metrics)),
                       int(sizeof(rpd_bilateral_config)), int(sizeof(rpd_atrous_config))};
    for (int k = 0; k < n && k < 14; ++k) out[k] = s[k];
    return RPD_OK;
}

int rpd_job_bytes(const rpd_render_config* cfg, uint64_t* bytes) {
    if (!cfg || !bytes) return fail(RPD_ERR_INVALID_ARG, "null argument");
    const int st = validate(*cfg);
    if (st) return st;
    *bytes = job_bytes(*cfg, 0, false, gather_overlap_applies(*cfg));
    return RPD_OK;
}

int rpd_ctx_memory(rpd_context* ctx, uint64_t* in_use, uint64_t* budget, uint64_t* device_free,
                   uint64_t* device_total) {
    if (!ctx) return fail(RPD_ERR_INVALID_ARG, "null context");
    if (in_use) *in_use = ctx->in_use.load();
    if (budget) *budget = ctx->budget;
    if (device_free || device_total) {
        size_t fr = 0, tot = 0;
        HIP_TRY(hipSetDevice(ctx->device));
        HIP_TRY(hipMemGetInfo(&fr, &tot));
        if (device_free) *device_free = fr;
        if (device_total) *device_total = tot;
    }
    return RPD_OK;
}

int rpd_job_grid_dims(rpd_job* j, int dims[4]) {
    if (!j || !dims) return fail(RPD_ERR_INVALID_ARG, "null argument");
    dims[0] = j->geom.nx; dims[1] = j->geom.ny; dims[2] = j->geom.step; dims[3] = j->geom.support;
    return RPD_OK;
}

int rpd_job_stage_ms(rpd_job* j, double* ms, int n) {
    if (!j || !ms) return fail(RPD_ERR_INVALID_ARG, "null argument");
    if (j->worker.joinable()) j->worker.join();
    HIP_TRY(hipStreamSynchronize(j->stream));
    collect_spans(j);
    double total = 0;
    for (int k = 0; k < 6; ++k) total += j->stage_ms[k];
    j->stage_ms[6] = total;
    for (int k = 0; k < n && k < 11; ++k) ms[k] = j->stage_ms[k];
    return RPD_OK;
}

// ---- image-quality metrics ---------------------------------------------------
// Pass A on the device, its fields to the host (means and log-luminance peak
// set the constants of pass B), pass B, then the closing formulas of
// hdr_metrics (shared/tools/eval_tools.cpp:236-279) and
// compute_extended_metrics (eval_tools.cpp:145, 172-174, 191, 212).
// The caller holds ctx->metrics_mu with the context's device current.
// Partial sums of both passes for a W x H image, in doubles.
static size_t metrics_part_doubles(int W, int H) {
    return size_t(std::max(metrics_a_blocks(W, H) * MA_FIELDS, metrics_b_blocks(W, H) * MB_FIELDS));
}

// Grows the context's partial and output buffers for a W x H image.
static int metrics_buffers(rpd_context* ctx, int W, int H) {
    const size_t need = metrics_part_doubles(W, H);
    if (need > ctx->m_part_cap) {
        if (ctx->m_part) HIP_TRY(dev_free(ctx->m_part));
        ctx->m_part = nullptr;
        ctx->m_part_cap = 0;
        HIP_TRY(dev_malloc(&ctx->m_part, need * sizeof(double)));
        ctx->m_part_cap = need;
    }
    if (!ctx->m_out) HIP_TRY(dev_malloc(&ctx->m_out, NOALLOC_METRICS_OUT * sizeof(double)));
    return RPD_OK;
}

// `part` holds metrics_part_doubles(W, H) doubles and `dout` 32; every
// partial the reductions read is written by the pass before, so neither needs
// clearing between calls.
static int metrics_run(const MetricImage& ref, const MetricImage& tst, int W, int H, double* part, double* dout,
                       hipStream_t st, rpd_image_metrics* out) {
    double a[MA_FIELDS], b[MB_FIELDS];
    launch_metrics_a(ref, tst, W, H, part, dout, st);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(a, dout, sizeof(a), hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    const double pixels = double(W) * double(H);
    MetricsPassB pb;
    for (int c = 0; c < 3; ++c) { pb.mean_ref[c] = a[MA_MEAN_REF + c] / pixels; pb.mean_tst[c] = a[MA_MEAN_TST + c] / pixels; }
    const double log_peak = std::max(1e-6, a[MA_LOG_PEAK]);
    pb.c1 = std::pow(0.01 * log_peak, 2.0);
    pb.c2 = std::pow(0.03 * log_peak, 2.0);
    launch_metrics_b(ref, tst, W, H, pb, part, dout, st);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(b, dout, sizeof(b), hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));

    const double peak = a[MA_PEAK] > 0.0 ? a[MA_PEAK] : 1.0;
    const double C1 = (0.01 * peak) * (0.01 * peak), C2 = (0.03 * peak) * (0.03 * peak);
    for (int c = 0; c < 3; ++c) {
        const double vr = b[MB_VAR_REF + c] / pixels, vt = b[MB_VAR_TST + c] / pixels, cov = b[MB_COV + c] / pixels;
        const double mr = pb.mean_ref[c], mt = pb.mean_tst[c];
        const double num = (2.0 * mr * mt + C1) * (2.0 * cov + C2);
        const double den = (mr * mr + mt * mt + C1) * (vr + vt + C2);
        out->ssim[c] = den > 0.0 ? num / den : 0.0;
    }
    out->ssim_mean = (out->ssim[0] + out->ssim[1] + out->ssim[2]) / 3.0;
    const double rmse = std::sqrt(a[MA_SUM_SQ] / (pixels * 3.0));
    out->mae = a[MA_SUM_ABS] / (pixels * 3.0);
    out->rmse = rmse;
    out->psnr = rmse > 0.0 ? 20.0 * std::log10(peak / rmse) : INFINITY;
    out->peak = peak;
    out->nrmse = std::sqrt(a[MA_SUM_SQ] / std::max(1e-18, a[MA_REF_ENERGY]));
    out->log_ssim11 = b[MB_SSIM] / pixels;
    out->log_gradient_nrmse = std::sqrt(a[MA_GRAD_ERR] / std::max(1e-18, a[MA_GRAD_REF]));
    out->relmse = a[MA_RELMSE] / (pixels * 3.0);
    return RPD_OK;
}

static int metrics_check_dims(int width, int height) {
    if (width <= 0 || height <= 0) return fail(RPD_ERR_INVALID_ARG, "image dimensions must be positive");
    if ((height + 15) / 16 > 65535) return fail(RPD_ERR_INVALID_ARG, "image taller than 1048560 rows");
    return RPD_OK;
}

// Grows the context's image staging area to `floats` and returns it.
static int metrics_stage(rpd_context* ctx, size_t floats, float** p) {
    if (floats > ctx->m_img_cap) {
        if (ctx->m_img) HIP_TRY(dev_free(ctx->m_img));
        ctx->m_img = nullptr;
        ctx->m_img_cap = 0;
        HIP_TRY(dev_malloc(&ctx->m_img, floats * sizeof(float)));
        ctx->m_img_cap = floats;
    }
    *p = ctx->m_img;
    return RPD_OK;
}

int rpd_image_metrics_compute(rpd_context* ctx, const float* ref, const float* test, int width, int height,
                              int on_device, rpd_image_metrics* out) {
    if (!ctx || !ref || !test || !out) return fail(RPD_ERR_INVALID_ARG, "null argument");
    int st = metrics_check_dims(width, height);
    if (st) return st;
    std::lock_guard<std::mutex> lk(ctx->metrics_mu);
    HIP_TRY(hipSetDevice(ctx->device));
    const size_t n3 = size_t(width) * size_t(height) * 3;
    const float* dr = ref;
    const float* dt = test;
    if (!on_device) {
        float* stage = nullptr;
        if ((st = metrics_stage(ctx, 2 * n3, &stage))) return st;
        HIP_TRY(hipMemcpyAsync(stage, ref, n3 * sizeof(float), hipMemcpyHostToDevice, ctx->m_stream));
        HIP_TRY(hipMemcpyAsync(stage + n3, test, n3 * sizeof(float), hipMemcpyHostToDevice, ctx->m_stream));
        dr = stage;
        dt = stage + n3;
    }
    if ((st = metrics_buffers(ctx, width, height))) return st;
    return metrics_run(MetricImage{dr, 1, 3}, MetricImage{dt, 1, 3}, width, height, ctx->m_part, ctx->m_out,
                       ctx->m_stream, out);
}

int rpd_image_metrics_compute_noalloc(rpd_context* ctx, const float* ref, const float* test, int width, int height,
                                      rpd_image_metrics* out) {
    if (!ctx || !ref || !test || !out) return fail(RPD_ERR_INVALID_ARG, "null argument");
    int st = metrics_check_dims(width, height);
    if (st) return st;
    if (metrics_b_blocks(width, height) > NOALLOC_METRICS_MAX_TILES ||
        metrics_part_doubles(width, height) > NOALLOC_METRICS_PARTS)
        return fail(RPD_ERR_INVALID_ARG, "image too large for rpd_image_metrics_compute_noalloc: at most 65536 "
                                         "16x16 tiles (4096 x 4096 pixels)");
    if (!ctx->scratch || !ctx->m_stream) return fail(RPD_ERR_STATE, "context has no metrics scratch");
    std::lock_guard<std::mutex> lk(ctx->metrics_mu);
    HIP_TRY(hipSetDevice(ctx->device));
    return metrics_run(MetricImage{ref, 1, 3}, MetricImage{test, 1, 3}, width, height, ctx->scratch,
                       ctx->scratch + NOALLOC_METRICS_PARTS, ctx->m_stream, out);
}

int rpd_job_image_metrics(rpd_job* j, int buffer_id, const float* ref, int ref_on_device, rpd_image_metrics* out) {
    if (!j || !ref || !out) return fail(RPD_ERR_INVALID_ARG, "null argument");
    const float* planes = nullptr;
    switch (buffer_id) {
        case RPD_BUF_RADIANCE:
            if (!j->finalized) return fail(RPD_ERR_STATE, "the radiance buffer holds sums until the job is finalized");
            planes = j->out_rad;
            break;
        case RPD_BUF_RAW_RADIANCE:
            if (!j->finalized) return fail(RPD_ERR_STATE, "the raw radiance is written when the job is finalized");
            planes = j->out_raw;
            break;
        case RPD_BUF_DENOISED: planes = j->den; break;
        default: return fail(RPD_ERR_INVALID_ARG, "metrics are defined for the radiance, raw radiance and denoised buffers");
    }
    if (j->worker.joinable()) j->worker.join();
    rpd_context* ctx = j->ctx;
    std::lock_guard<std::mutex> lk(ctx->metrics_mu);
    HIP_TRY(hipSetDevice(ctx->device));
    const size_t n3 = size_t(j->N) * 3;
    const float* dr = ref;
    if (!ref_on_device) {
        float* stage = nullptr;
        const int st = metrics_stage(ctx, n3, &stage);
        if (st) return st;
        HIP_TRY(hipMemcpyAsync(stage, ref, n3 * sizeof(float), hipMemcpyHostToDevice, j->stream));
        dr = stage;
    }
    const int st = metrics_buffers(ctx, j->W, j->H);
    if (st) return st;
    return metrics_run(MetricImage{dr, 1, 3}, MetricImage{planes, int64_t(j->N), 1}, j->W, j->H, ctx->m_part,
                       ctx->m_out, j->stream, out);
}

// ---- comparison denoisers ---------------------------------------------------
void rpd_bilateral_config_defaults(rpd_bilateral_config* c) {
    if (!c) return;
    c->sigma_c = 1.0f; c->sigma_n = 1.0f; c->sigma_p = 1.0f; c->sigma_a = 1.0f;  // joint_bilateral.cpp:172-181
}
void rpd_atrous_config_defaults(rpd_atrous_config* c) {
    if (!c) return;
    c->sigma_c = 0.5f; c->sigma_n = 0.5f; c->sigma_a = 0.1f; c->iterations = 5;  // atrous_wavelet.cpp:220-230
}

// The context's filter scratch of at least `floats` floats, on its stream.
static int filters_scratch(rpd_context* ctx, size_t floats, float** p) {
    if (!ctx->f_stream) HIP_TRY(stream_create(&ctx->f_stream));
    if (floats > ctx->f_cap) {
        if (ctx->f_buf) HIP_TRY(dev_free(ctx->f_buf));
        ctx->f_buf = nullptr;
        ctx->f_cap = 0;
        HIP_TRY(dev_malloc(&ctx->f_buf, floats * sizeof(float)));
        ctx->f_cap = floats;
    }
    *p = ctx->f_buf;
    return RPD_OK;
}

static int filter_dims(int width, int height) {
    if (width <= 0 || height <= 0) return fail(RPD_ERR_INVALID_ARG, "image dimensions must be positive");
    if ((height + 15) / 16 > 65535) return fail(RPD_ERR_INVALID_ARG, "image taller than 1048560 rows");
    return RPD_OK;
}

int rpd_joint_bilateral(rpd_context* ctx, const float* raw, const float* normal, const float* albedo, int width,
                        int height, const rpd_bilateral_config* cfg, float* gaussian, float* bilateral, float* joint,
                        int on_device) {
    if (!ctx || !raw || !normal || !albedo || !cfg) return fail(RPD_ERR_INVALID_ARG, "null argument");
    if (!gaussian && !bilateral && !joint) return fail(RPD_ERR_INVALID_ARG, "no output requested");
    int st = filter_dims(width, height);
    if (st) return st;
    if (!(cfg->sigma_c > 0.0f) || !(cfg->sigma_n > 0.0f) || !(cfg->sigma_p > 0.0f) || !(cfg->sigma_a > 0.0f))
        return fail(RPD_ERR_INVALID_ARG, "All sigma values must be positive");  // joint_bilateral.cpp:283-287
    if (bilateral_radius(cfg->sigma_p) > 96) return fail(RPD_ERR_INVALID_ARG, "sigma_p above 32 (window radius beyond 96)");
    for (const float* o : {static_cast<const float*>(gaussian), static_cast<const float*>(bilateral),
                           static_cast<const float*>(joint)})
        if (o && (o == raw || o == normal || o == albedo)) return fail(RPD_ERR_INVALID_ARG, "outputs must not alias an input");
    std::lock_guard<std::mutex> lk(ctx->filters_mu);
    HIP_TRY(hipSetDevice(ctx->device));
    const size_t n3 = size_t(width) * size_t(height) * 3;
    float* buf = nullptr;
    if ((st = filters_scratch(ctx, on_device ? 0 : 6 * n3, &buf))) return st;
    hipStream_t s = ctx->f_stream;
    BilateralParams bp;
    bp.width = width; bp.height = height;
    bp.sigma_c = cfg->sigma_c; bp.sigma_n = cfg->sigma_n; bp.sigma_p = cfg->sigma_p; bp.sigma_a = cfg->sigma_a;
    bp.radius = bilateral_radius(cfg->sigma_p);
    if (on_device) {
        bp.raw = raw; bp.normal = normal; bp.albedo = albedo;
        bp.gaussian = gaussian; bp.bilateral = bilateral; bp.joint = joint;
    } else {
        HIP_TRY(hipMemcpyAsync(buf, raw, n3 * sizeof(float), hipMemcpyHostToDevice, s));
        HIP_TRY(hipMemcpyAsync(buf + n3, normal, n3 * sizeof(float), hipMemcpyHostToDevice, s));
        HIP_TRY(hipMemcpyAsync(buf + 2 * n3, albedo, n3 * sizeof(float), hipMemcpyHostToDevice, s));
        bp.raw = buf; bp.normal = buf + n3; bp.albedo = buf + 2 * n3;
        bp.gaussian = gaussian ? buf + 3 * n3 : nullptr;
        bp.bilateral = bilateral ? buf + 4 * n3 : nullptr;
        bp.joint = joint ? buf + 5 * n3 : nullptr;
    }
    launch_joint_bilateral(bp, s);
    HIP_TRY(hipGetLastError());
    if (!on_device) {
        if (gaussian) HIP_TRY(hipMemcpyAsync(gaussian, buf + 3 * n3, n3 * sizeof(float), hipMemcpyDeviceToHost, s));
        if (bilateral) HIP_TRY(hipMemcpyAsync(bilateral, buf + 4 * n3, n3 * sizeof(float), hipMemcpyDeviceToHost, s));
        if (joint) HIP_TRY(hipMemcpyAsync(joint, buf + 5 * n3, n3 * sizeof(float), hipMemcpyDeviceToHost, s));
    }
    HIP_TRY(hipStreamSynchronize(s));
    return RPD_OK;
}

int rpd_atrous(rpd_context* ctx, const float* raw, const float* normal, const float* albedo, int width, int height,
               const rpd_atrous_config* cfg, float* out, int on_device) {
    if (!ctx || !raw || !normal || !albedo || !cfg || !out) return fail(RPD_ERR_INVALID_ARG, "null argument");
    int st = filter_dims(width, height);
    if (st) return st;
    if (!(cfg->sigma_c > 0.0f) || !(cfg->sigma_n > 0.0f) || !(cfg->sigma_a > 0.0f))
        return fail(RPD_ERR_INVALID_ARG, "All sigma values must be positive");  // atrous_wavelet.cpp:363-369
    if (cfg->iterations < 1) return fail(RPD_ERR_INVALID_ARG, "Number of iterations must be at least 1");
    if (cfg->iterations > 30) return fail(RPD_ERR_INVALID_ARG, "more than 30 iterations (step 2^iteration overflows)");
    if (out == raw || out == normal || out == albedo) return fail(RPD_ERR_INVALID_ARG, "out must not alias an input");
    std::lock_guard<std::mutex> lk(ctx->filters_mu);
    HIP_TRY(hipSetDevice(ctx->device));
    const size_t n = size_t(width) * size_t(height), n3 = 3 * n;
    // scratch: ping-pong image, unit normals, luminance [, staged raw, normal, albedo, out]
    float* buf = nullptr;
    if ((st = filters_scratch(ctx, 7 * n + (on_device ? 0 : 4 * n3), &buf))) return st;
    hipStream_t s = ctx->f_stream;
    float* tmp = buf;
    float* unit = buf + n3;
    float* lum = buf + 2 * n3;
    const float* d_raw = raw;
    const float* d_nrm = normal;
    const float* d_alb = albedo;
    float* d_out = out;
    if (!on_device) {
        float* stage = buf + 7 * n;
        HIP_TRY(hipMemcpyAsync(stage, raw, n3 * sizeof(float), hipMemcpyHostToDevice, s));
        HIP_TRY(hipMemcpyAsync(stage + n3, normal, n3 * sizeof(float), hipMemcpyHostToDevice, s));
        HIP_TRY(hipMemcpyAsync(stage + 2 * n3, albedo, n3 * sizeof(float), hipMemcpyHostToDevice, s));
        d_raw = stage; d_nrm = stage + n3; d_alb = stage + 2 * n3; d_out = stage + 3 * n3;
    }
    launch_atrous_prepare(d_nrm, int64_t(n), unit, s);
    // pass k writes `out` when iterations - 1 - k is even, so the last pass ends there
    const float* src = d_raw;
    for (int k = 0; k < cfg->iterations; ++k) {
        float* dst = ((cfg->iterations - 1 - k) % 2 == 0) ? d_out : tmp;
        AtrousPass ap;
        ap.in = src; ap.out = dst; ap.unit = unit; ap.albedo = d_alb; ap.lum = lum;
        ap.width = width; ap.height = height; ap.step = 1 << k;
        ap.sigma_c = cfg->sigma_c; ap.sigma_n = cfg->sigma_n; ap.sigma_a = cfg->sigma_a;
        launch_atrous_pass(ap, s);
        src = dst;
    }
    HIP_TRY(hipGetLastError());
    if (!on_device) HIP_TRY(hipMemcpyAsync(out, d_out, n3 * sizeof(float), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    return RPD_OK;
}

int rpd_job_counters(rpd_job* j, uint64_t* out, int n) {
    if (!j || !out) return fail(RPD_ERR_INVALID_ARG, "null argument");
    if (j->worker.joinable()) j->worker.join();
    HIP_TRY(hipStreamSynchronize(j->stream));
    unsigned long long q[2] = {0, 0};
    HIP_TRY(sync_memcpy(q, j->query_count, sizeof(q), hipMemcpyDeviceToHost));
    const uint64_t v[4] = {q[0], q[1], j->scene ? uint64_t(j->scene->n_tris) : 0,
                           j->scene ? uint64_t(j->scene->n_spheres) : 0};
    for (int k = 0; k < n && k < 4; ++k) out[k] = v[k];
    return RPD_OK;
}

int rpd_job_buffer_size(rpd_job* j, int id, size_t* bytes) {
    if (!j || !bytes) return fail(RPD_ERR_INVALID_ARG, "null argument");
    const size_t N = j->N, nn = size_t(j->geom.nx) * j->geom.ny;
    switch (id) {
        case RPD_BUF_RADIANCE: case RPD_BUF_RAW_RADIANCE: case RPD_BUF_ALBEDO: case RPD_BUF_NORMAL:
        case RPD_BUF_WORLD_POS: case RPD_BUF_DENOISED: *bytes = 12 * N; return RPD_OK;
        case RPD_BUF_STATS: *bytes = 48 * N; return RPD_OK;
        case RPD_BUF_SUMS: *bytes = 52 * N; return RPD_OK;
        case RPD_BUF_FEATURES: *bytes = 28 * N; return RPD_OK;
        case RPD_BUF_GRID: *bytes = 8 * GRID_NODE_DOUBLES * nn; return RPD_OK;
        case RPD_BUF_TILES: *bytes = 4 * SF_COUNT * nn; return RPD_OK;
        case RPD_BUF_RECORDS: {
            int tot[2] = {0, 0};
            HIP_TRY(hipStreamSynchronize(j->stream));
            HIP_TRY(sync_memcpy(tot, j->totals, sizeof(tot), hipMemcpyDeviceToHost));
            *bytes = size_t(tot[1] > 0 ? tot[0] : 0) * 32 * sizeof(float);
            return RPD_OK;
        }
        default:
            if (id >= RPD_BUF_DEBUG_NEIGHBORS && id <= RPD_BUF_DEBUG_VARIANCE_SOURCE) {
                if (!j->dbg) return fail(RPD_ERR_STATE, "debug buffers are kept only for jobs created after "
                                                        "rpd_ctx_set_debug_outputs(ctx, 1)");
                *bytes = 4 * N;
                return RPD_OK;
            }
            if ((id >= RPD_BUF_DEPTH && id <= RPD_BUF_SAMPLE_FRACTION) || (id >= RPD_BUF_SENS_ALL && id <= RPD_BUF_SENS_RR) ||
                (id >= RPD_BUF_SAMPLE_COUNT && id <= RPD_BUF_PASS_COUNTS)) {
                *bytes = 4 * N;
                return RPD_OK;
            }
            return fail(RPD_ERR_INVALID_ARG, "unknown buffer id");
    }
}

int rpd_job_read(rpd_job* j, int id, void* dst, size_t bytes) {
    if (!j || !dst) return fail(RPD_ERR_INVALID_ARG, "null argument");
    size_t need = 0;
    int st = rpd_job_buffer_size(j, id, &need);
    if (st) return st;
    if (bytes < need) return fail(RPD_ERR_INVALID_ARG, "destination buffer too small");
    if (j->worker.joinable()) j->worker.join();
    HIP_TRY(hipStreamSynchronize(j->stream));
    const size_t N = j->N;
    float* f = static_cast<float*>(dst);
    auto scalar = [&](const float* d) -> int {
        HIP_TRY(sync_memcpy(dst, d, 4 * N, hipMemcpyDeviceToHost));
        return RPD_OK;
    };
    switch (id) {
        case RPD_BUF_RADIANCE:
            return copy_planes_to_interleaved(j, j->finalized ? j->out_rad : j->pb.radiance, 3, f);
        case RPD_BUF_RAW_RADIANCE: return copy_planes_to_interleaved(j, j->out_raw, 3, f);
        case RPD_BUF_ALBEDO: return copy_planes_to_interleaved(j, j->finalized ? j->out_alb : j->pb.albedo, 3, f);
        case RPD_BUF_NORMAL: return copy_planes_to_interleaved(j, j->finalized ? j->out_nrm : j->pb.normal, 3, f);
        case RPD_BUF_WORLD_POS: return copy_planes_to_interleaved(j, j->finalized ? j->out_pos : j->pb.world_pos, 3, f);
        case RPD_BUF_DENOISED: return copy_planes_to_interleaved(j, j->den, 3, f);
        case RPD_BUF_DEPTH: return scalar(j->finalized ? j->out_dep : j->pb.depth);
        case RPD_BUF_UNCERTAINTY: return scalar(j->unc);
        case RPD_BUF_ALPHA: return scalar(j->alp);
        case RPD_BUF_VARIANCE_DENOISED: return scalar(j->vden);
        case RPD_BUF_VAR_TOTAL: return scalar(j->vtot);
        case RPD_BUF_VAR_EFF: return scalar(j->veff);
        case RPD_BUF_VAR_RATIO: return scalar(j->vrat);
        case RPD_BUF_SENSITIVITY_CONFIDENCE: return scalar(j->sconf);
        case RPD_BUF_SENSITIVITY_GRADIENT: return scalar(j->sgrad);
        case RPD_BUF_ADAPTIVE_IMPORTANCE: return scalar(j->imp);
        case RPD_BUF_DEBUG_NEIGHBORS: return scalar(j->dbg);
        case RPD_BUF_DEBUG_WEIGHT_SUM: return scalar(j->dbg + N);
        case RPD_BUF_DEBUG_NEFF: return scalar(j->dbg + 2 * N);
        case RPD_BUF_DEBUG_VARIANCE_SOURCE: return scalar(j->dbg + 3 * N);
        case RPD_BUF_SAMPLE_COUNT: HIP_TRY(sync_memcpy(dst, j->pb.sample_count, 4 * N, hipMemcpyDeviceToHost)); return RPD_OK;
        case RPD_BUF_HIT_COUNT: HIP_TRY(sync_memcpy(dst, j->pb.hit_count, 4 * N, hipMemcpyDeviceToHost)); return RPD_OK;
        case RPD_BUF_PASS_COUNTS: HIP_TRY(sync_memcpy(dst, j->counts, 4 * N, hipMemcpyDeviceToHost)); return RPD_OK;
        case RPD_BUF_SAMPLE_FRACTION: {  // core/buffers.h:182-195
            std::vector<int> c(N);
            HIP_TRY(sync_memcpy(c.data(), j->pb.sample_count, 4 * N, hipMemcpyDeviceToHost));
            const int mx = std::max(1, N ? *std::max_element(c.begin(), c.end()) : 1);
            const float inv = 1.0f / float(mx);
            for (size_t i = 0; i < N; ++i) f[i] = float(std::max(0, c[i])) * inv;
            return RPD_OK;
        }
        case RPD_BUF_STATS: {
            std::vector<float> s(12 * N);
            HIP_TRY(sync_memcpy(s.data(), j->pb.lum_mean, 48 * N, hipMemcpyDeviceToHost));
            const int plane_of[12] = {0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11};
            for (size_t i = 0; i < N; ++i)
                for (int k = 0; k < 12; ++k) f[i * 12 + k] = s[size_t(plane_of[k]) * N + i];
            return RPD_OK;
        }
        case RPD_BUF_SUMS: {
            std::vector<float> s(13 * N);
            std::vector<int> ic(2 * N);
            HIP_TRY(sync_memcpy(s.data(), j->pb.radiance, 52 * N, hipMemcpyDeviceToHost));
            for (size_t i = 0; i < N; ++i)
                for (int k = 0; k < 13; ++k) f[i * 13 + k] = s[size_t(k) * N + i];
            return RPD_OK;
        }
        case RPD_BUF_FEATURES: {
            int rc = RPD_OK;
            ensure_features(j);
            std::vector<float> s(7 * N);
            HIP_TRY(hipMemcpyAsync(s.data(), j->feat.normal, 28 * N, hipMemcpyDeviceToHost, j->stream));
            HIP_TRY(hipStreamSynchronize(j->stream));
            for (size_t i = 0; i < N; ++i)
                for (int k = 0; k < 7; ++k) f[i * 7 + k] = s[size_t(k) * N + i];
            return rc;
        }
        case RPD_BUF_GRID:
            HIP_TRY(sync_memcpy(dst, j->grid, need, hipMemcpyDeviceToHost));
            return RPD_OK;
        case RPD_BUF_TILES:
            HIP_TRY(sync_memcpy(dst, j->tiles, need, hipMemcpyDeviceToHost));
            return RPD_OK;
        case RPD_BUF_RECORDS: {
            const size_t ns = need / (32 * sizeof(float));
            std::vector<float4> r[6];
            std::vector<float2> r6(ns);
            Records& R = j->rec;
            float4* srcs[6] = {R.r0, R.r1, R.r2, R.r3, R.r4, R.r5};
            for (int k = 0; k < 6; ++k) {
                r[k].resize(ns);
                if (ns) HIP_TRY(sync_memcpy(r[k].data(), srcs[k], ns * sizeof(float4), hipMemcpyDeviceToHost));
            }
            if (ns) HIP_TRY(sync_memcpy(r6.data(), R.r6, ns * sizeof(float2), hipMemcpyDeviceToHost));
            std::vector<int> offs(N), cnts(N);
            HIP_TRY(sync_memcpy(offs.data(), j->offsets, 4 * N, hipMemcpyDeviceToHost));
            HIP_TRY(sync_memcpy(cnts.data(), j->counts, 4 * N, hipMemcpyDeviceToHost));
            // 32 floats per sample in pixel-major sample order (same layout as the oracle's records)
            for (size_t p = 0; p < N; ++p) {
                for (int s = 0; s < cnts[p]; ++s) {
                    const size_t k = size_t(offs[p]) + s;
                    float* o = f + k * 32;
                    std::memset(o, 0, 32 * sizeof(float));
                    uint32_t fl;
                    std::memcpy(&fl, &r[0][k].w, 4);
                    const bool hit = fl & REC_HIT;
                    o[0] = float(p); o[1] = r[0][k].x; o[2] = r[0][k].y; o[3] = r[0][k].z;
                    o[4] = (fl & REC_FINITE) ? 1.0f : 0.0f; o[5] = hit ? 1.0f : 0.0f;
                    if (hit) {
                        o[6] = r[1][k].x; o[7] = r[1][k].y; o[8] = r[1][k].z; o[15] = r[1][k].w;
                        o[9] = r[2][k].x; o[10] = r[2][k].y; o[11] = r[2][k].z;
                        o[12] = r[3][k].x; o[13] = r[3][k].y; o[14] = r[3][k].z;
                    }
                    if (j->last_pass_capture) {  // dimensions that were never captured read as zero
                        if (fl & REC_PIXEL) { o[16] = r[4][k].x; o[17] = r[4][k].y; }
                        if (fl & REC_LENS) { o[18] = r[4][k].z; o[19] = r[4][k].w; }
                        if (fl & REC_BRDF) { o[20] = r[5][k].x; o[21] = r[5][k].y; }
                        if (fl & REC_LIGHT) { o[22] = r[5][k].z; o[23] = r[5][k].w; }
                        if (fl & REC_ENV) { o[24] = r6[k].x; o[25] = r6[k].y; }
                    }
                    const uint32_t li = (fl >> REC_LIGHT_SHIFT) & 255u;
                    o[26] = (fl & REC_LSEL) ? float(li) : -1.0f;
                    o[27] = (fl & REC_RR_SURVIVED) ? 1.0f : 0.0f;
                    o[28] = float(((fl & REC_PIXEL) ? 1 : 0) | ((fl & REC_LENS) ? 2 : 0) | ((fl & REC_LSEL) ? 4 : 0) |
                                  ((fl & REC_BRDF) ? 8 : 0) | ((fl & REC_LIGHT) ? 16 : 0) | ((fl & REC_ENV) ? 32 : 0) |
                                  ((fl & REC_RR) ? 64 : 0));
                }
            }
            return RPD_OK;
        }
        default: break;
    }
    if (id >= RPD_BUF_SENS_ALL && id <= RPD_BUF_SENS_RR) {
        const int field[9] = {SF_ALL, SF_PIXEL, SF_BRDF, SF_LENS, SF_LIGHT, SF_LIGHT_UV, SF_LIGHT_SELECT, SF_ENV, SF_RR};
        const size_t nn = size_t(j->geom.nx) * j->geom.ny;
        std::vector<float> t(nn);
        HIP_TRY(sync_memcpy(t.data(), j->tiles + size_t(field[id - RPD_BUF_SENS_ALL]) * nn, 4 * nn, hipMemcpyDeviceToHost));
        expand_tiles(j, t.data(), f);
        return RPD_OK;
    }
    return fail(RPD_ERR_INVALID_ARG, "buffer is not readable");
}

int rpd_job_write(rpd_job* j, int id, const void* src, size_t bytes) {
    if (!j || !src) return fail(RPD_ERR_INVALID_ARG, "null argument");
    size_t need = 0;
    int st = rpd_job_buffer_size(j, id, &need);
    if (st) return st;
    if (bytes < need) return fail(RPD_ERR_INVALID_ARG, "source buffer too small");
    HIP_TRY(hipStreamSynchronize(j->stream));
    const size_t N = j->N;
    const float* f = static_cast<const float*>(src);
    j->features_valid = false;
    switch (id) {
        case RPD_BUF_STATS: {
            std::vector<float> s(12 * N);
            for (size_t i = 0; i < N; ++i)
                for (int k = 0; k < 12; ++k) s[size_t(k) * N + i] = f[i * 12 + k];
            HIP_TRY(sync_memcpy(j->pb.lum_mean, s.data(), 48 * N, hipMemcpyHostToDevice));
            return RPD_OK;
        }
        case RPD_BUF_SUMS: {
            std::vector<float> s(13 * N);
            for (size_t i = 0; i < N; ++i)
                for (int k = 0; k < 13; ++k) s[size_t(k) * N + i] = f[i * 13 + k];
            HIP_TRY(sync_memcpy(j->pb.radiance, s.data(), 52 * N, hipMemcpyHostToDevice));
            return RPD_OK;
        }
        case RPD_BUF_GRID: HIP_TRY(sync_memcpy(j->grid, src, need, hipMemcpyHostToDevice)); return RPD_OK;
        case RPD_BUF_TILES: HIP_TRY(sync_memcpy(j->tiles, src, need, hipMemcpyHostToDevice)); return RPD_OK;
        case RPD_BUF_HIT_COUNT: HIP_TRY(sync_memcpy(j->pb.hit_count, src, 4 * N, hipMemcpyHostToDevice)); return RPD_OK;
        case RPD_BUF_SAMPLE_COUNT: HIP_TRY(sync_memcpy(j->pb.sample_count, src, 4 * N, hipMemcpyHostToDevice)); return RPD_OK;
        case RPD_BUF_SENSITIVITY_GRADIENT: HIP_TRY(sync_memcpy(j->sgrad, src, 4 * N, hipMemcpyHostToDevice)); return RPD_OK;
        case RPD_BUF_VARIANCE_DENOISED: HIP_TRY(sync_memcpy(j->vden, src, 4 * N, hipMemcpyHostToDevice)); return RPD_OK;
        default: return fail(RPD_ERR_INVALID_ARG, "buffer is not writable");
    }
}

}  // extern "C"
