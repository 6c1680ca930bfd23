/*
 * rpd_engine.h - C ABI of the MI355X path-trace -> RPF grid -> StatMC ->
 * adaptive-pass engine.
 *
 * The reference renderer (andersonliu1/rpd-statmc-denoiser) has no library
 * interface: its hot path is the function seam
 *     render(const RenderConfig&, const Camera&)      path_tracer/main.cpp:1543
 * over process globals (main.cpp:38-43), followed by
 *     output_buffers(const FrameBuffers&, ...)        path_tracer/core/buffers.h:197
 * Each entry point below names the reference function it stands in for.
 *
 * Conventions: every function returns an rpd_status (0 = ok, negative = error)
 * and never throws; rpd_last_error() returns a thread-local message for the
 * last failure on the calling thread. All pointers are host pointers owned by
 * the caller; the engine copies what it needs. No GPU-runtime types appear.
 */
#ifndef RPD_ENGINE_H
#define RPD_ENGINE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum rpd_status {
    RPD_OK = 0,
    RPD_ERR_INVALID_ARG = -1,
    RPD_ERR_OUT_OF_MEMORY = -2, /* device allocation failed or admission refused */
    RPD_ERR_DEVICE = -3,        /* HIP runtime error */
    RPD_ERR_STATE = -4,         /* call not valid in the object's current state */
    RPD_ERR_UNSUPPORTED = -5
} rpd_status;

typedef struct rpd_context rpd_context;
typedef struct rpd_scene rpd_scene;
typedef struct rpd_job rpd_job;

/* ---- scene description (path_tracer/core/{triangle,sphere,material,light}.h) ---- */
enum { RPD_MAT_LAMBERTIAN = 0, RPD_MAT_MICROFACET = 1, RPD_MAT_MIRROR = 2 };
enum { RPD_MICROFACET_GGX = 0, RPD_MICROFACET_BECKMANN = 1 };
enum { RPD_LIGHT_RECT = 0, RPD_LIGHT_AREA = 1, RPD_LIGHT_POINT = 2, RPD_LIGHT_DIRECTIONAL = 3 };

typedef struct rpd_triangle { /* core/triangle.h:6-11 */
    float v0[3], v1[3], v2[3];
    float normal[3];          /* shading normal as stored by the scene (may be flipped) */
    float emission[3];
    int material_id;
    int light_id;             /* light this emitter is linked to, -1 if none (core/scene.h:41-43) */
} rpd_triangle;

typedef struct rpd_sphere { /* core/sphere.h:6-11 */
    float center[3];
    float radius;
    float emission[3];
    int material_id;
    int light_id;
} rpd_sphere;

typedef struct rpd_material { /* core/material.h:9-240 */
    int kind;                 /* RPD_MAT_* */
    float albedo[3];
    float roughness;          /* microfacet only */
    float n1[3], n2[3];       /* microfacet Schlick IORs */
    int distribution;         /* RPD_MICROFACET_* */
} rpd_material;

typedef struct rpd_light { /* core/light.h:16-150 */
    int kind;                 /* RPD_LIGHT_* */
    float p0[3];              /* rect: corner; area: v0; point: position; directional: direction */
    float p1[3];              /* rect: width edge; area: v1 */
    float p2[3];              /* rect: length edge; area: v2 */
    float normal[3];          /* area: triangle normal (rect computes its own) */
    float emission[3];        /* radiance, or intensity for point/directional */
} rpd_light;

/* ---- render configuration: POD mirror of RenderConfig (core/config.h:15-49) ---- */
typedef struct rpd_render_config {
    int image_width, image_height, samples_per_pixel;
    int use_statmc;
    int rpf_tile_size;
    int color_window_radius;
    float color_normal_threshold, color_depth_threshold, color_compat_alpha, color_sigma_max;
    int var_window_radius;
    float var_normal_threshold, var_depth_threshold, var_compat_sigma, var_shrinkage_k;
    int var_iterations;
    int adaptive_base_samples, adaptive_spp;
    float adaptive_sigma_max;
    int adaptive_passes, adaptive_importance_smoothing_radius;
    float rpf_shrinkage_scale, rpf_confidence_samples;
    /* Extension (no counterpart in the reference, whose paths end only by
       Russian roulette): at most max_depth shading points per path; the last
       one draws its BRDF sample but traces no BRDF ray. 0 = unbounded. */
    int max_depth;
} rpd_render_config;

/* Fills `cfg` with the reference defaults (core/config.h:27-48). */
void rpd_render_config_defaults(rpd_render_config* cfg);

typedef struct rpd_camera { /* inputs of Camera::init, core/camera.h:31-54 */
    float position[3], direction[3], world_up[3];
    float fov_degrees;
    float focus_distance;
    float aperture;           /* lens diameter; 0 = pinhole */
} rpd_camera;

/* ---- context ---- */
/* One context per device; thread-safe for job submission and polling. */
int rpd_ctx_create(int device, rpd_context** out);
int rpd_ctx_destroy(rpd_context* ctx);
const char* rpd_last_error(void);
/* Device memory the context may hand to jobs (admission budget), bytes. */
int rpd_ctx_set_memory_budget(rpd_context* ctx, uint64_t bytes);
/* Extension (the reference allocates its buffers once per process): a job's
   device buffers live in one slab, sized in classes of eight per doubling.
   Slabs of released jobs are kept, up to `bytes` (default: a quarter of the
   device's memory, at most 64 GiB; the oldest are freed first), and a later job takes the smallest kept slab that holds
   it and is at most half as large again; kept slabs do not count as held by
   jobs. A device allocation that fails empties the cache and is tried again.
   0 frees and disables it. `cached` (may be NULL) receives the bytes kept. */
int rpd_ctx_set_buffer_cache(rpd_context* ctx, uint64_t bytes, uint64_t* cached);
/* Extension: a job takes a kept slab larger than it needs only while the
   bytes held by jobs stay within `bytes` (0, the default: no limit). The
   serving loop sets its budget here, so that admission by footprint holds. */
int rpd_ctx_set_held_limit(rpd_context* ctx, uint64_t bytes);
/* Extension: before each pass the engine writes the first
   min(624, words_per_sample * count + margin) words (rounded up to 4) of a
   pixel's MT19937 state after its first twist, count being the pixel's samples
   in the pass; a pixel that draws more writes the rest during the trace.
   Draws are the same for any setting (defaults 48 and 64). */
int rpd_ctx_set_tape_prefix(rpd_context* ctx, int words_per_sample, int margin);
/* How scenes created next build their BVH (only scenes too large for the flat
   loop get one; the reference builds a binary median-split BVH on the host,
   core/ray_tracer.h:49-226): RPD_BVH_HOST_SAH (default) bins a surface-area
   heuristic on the host; RPD_BVH_DEVICE_LBVH builds a linear BVH from Morton
   codes on the device in milliseconds, trees slower to traverse, and falls back to
   the host build when its tree is too deep for the traversal stack. Hits do
   not depend on the tree. */
enum { RPD_BVH_HOST_SAH = 0, RPD_BVH_DEVICE_LBVH = 1 };
int rpd_ctx_set_bvh_builder(rpd_context* ctx, int builder);

/* ---- scenes: scenes::load_scene (core/scene_loader.h:11-13) + BVH build (main.cpp:1855) ---- */
int rpd_scene_create(rpd_context* ctx, const rpd_triangle* tris, int n_tris, const rpd_sphere* spheres,
                     int n_spheres, const rpd_material* materials, int n_materials, const rpd_light* lights,
                     int n_lights, const float environment[3], rpd_scene** out);
/* The reference's built-in scenes by name (scenes/scene_registry.h:17-53):
   cornell_box, cornell_spheres, veach_mis, material_showcase, caustics,
   debug, and the mesh scenes bunny, bunny_dof and dragon, which read
   resources/models/{bunny,dragon}.obj relative to the working directory as
   the reference does. ctx NULL keeps the scene on the host (rpd_scene_query). */
int rpd_scene_builtin(rpd_context* ctx, const char* name, rpd_scene** out);
/* The same scene as arrays. Pass NULL arrays to get the counts; otherwise
   each n_* holds the array's capacity on entry and the count on return. */
int rpd_builtin_scene_arrays(const char* name, rpd_triangle* tris, int* n_tris, rpd_sphere* spheres, int* n_spheres,
                             rpd_material* materials, int* n_materials, rpd_light* lights, int* n_lights,
                             float environment[3]);
/* Triangles of an OBJ mesh (core/obj_loader.h:19-65): every face's vertex
   positions (polygons split into a fan from their first vertex), normal from
   the winding, no emission, the given material. tris NULL: count only. */
int rpd_obj_load(const char* path, int material_id, rpd_triangle* tris, int* n_tris);
int rpd_scene_counts(const rpd_scene* scene, int* n_tris, int* n_spheres, int* n_materials, int* n_lights);
int rpd_scene_destroy(rpd_scene* scene);

/* ---- whole jobs: render() main.cpp:1543-1564 ---- */
/* Queues the full pipeline (base pass, adaptive passes, final StatMC) on the
   job's own stream and returns immediately. */
int rpd_job_submit(rpd_context* ctx, rpd_scene* scene, const rpd_render_config* cfg, const rpd_camera* cam,
                   uint32_t seed, rpd_job** out);
int rpd_job_poll(rpd_job* job, int* done);
int rpd_job_wait(rpd_job* job);
int rpd_job_release(rpd_job* job);
/* render() in two parts, for a caller that batches the second part of several
   jobs (the serving loop does). rpd_job_submit_passes queues the sampling part
   only: the base pass and the adaptive passes (main.cpp:1543-1547, 1451-1538);
   poll and wait then report that part. A job without StatMC is complete after
   it. rpd_render_finish_batch runs the denoise tail of every listed job whose
   passes have finished (main.cpp:1548-1564: sensitivity from the final grid,
   StatMC, variance denoise, outputs), the sensitivity and StatMC stages as one
   launch each over all the images; it returns with the outputs written,
   identical to rpd_job_submit's. gpu_ms (optional) receives the batch's device
   time. RPD_ERR_STATE for a job whose passes were not started this way or
   whose tail already ran. */
int rpd_job_submit_passes(rpd_context* ctx, rpd_scene* scene, const rpd_render_config* cfg, const rpd_camera* cam,
                          uint32_t seed, rpd_job** out);
int rpd_render_finish_batch(rpd_context* ctx, rpd_job* const* jobs, int n, double* gpu_ms);

/* ---- step-wise jobs: the reference's stage functions on one job's state ---- */
/* Allocates a job's device state (buffers.init + pixel_stats + rpf_grid.init,
   main.cpp:651-654) without queuing work. */
int rpd_job_create(rpd_context* ctx, rpd_scene* scene, const rpd_render_config* cfg, const rpd_camera* cam,
                   uint32_t seed, rpd_job** out);
/* One sampling pass (main.cpp:649-735 and 1451-1535). counts == NULL samples
   every pixel samples_per_pixel times; otherwise counts[i] times. statmc != 0
   seeds each pixel with its finite-sample count, folds Welford statistics and
   splats RPF samples; statmc == 0 is the uniform render (main.cpp:590-642). */
int rpd_job_run_pass(rpd_job* job, const int* counts, int statmc);
/* Adaptive pass: allocation from the current state, then the pass (main.cpp:1451-1538). */
int rpd_job_run_adaptive_pass(rpd_job* job);
int rpd_job_recompute_sensitivity(rpd_job* job);    /* main.cpp:741 */
int rpd_job_statmc_denoise(rpd_job* job);           /* main.cpp:957 */
int rpd_job_variance_denoise(rpd_job* job);         /* main.cpp:1166 */
int rpd_job_adaptive_counts(rpd_job* job, int* counts_out); /* main.cpp:1273 */
/* The part of compute_adaptive_sample_counts after the importance field is
   known (main.cpp:1373-1438): clipping, floors, systematic residual draw. */
int rpd_job_adaptive_counts_from_importance(rpd_job* job, const float* importance, int* counts_out);
int rpd_job_finalize(rpd_job* job);                 /* FrameBuffers::finalize, core/buffers.h:152 */
int rpd_job_sync(rpd_job* job);

/* ---- buffers (core/buffers.h:197-250 output set and internal state) ---- */
enum rpd_buffer_id {
    /* images: float RGB interleaved per pixel */
    RPD_BUF_RADIANCE = 0,          /* final image: denoised for StatMC, mean radiance otherwise */
    RPD_BUF_RAW_RADIANCE = 1,      /* StatMC only: mean radiance before reconstruction */
    RPD_BUF_ALBEDO = 2, RPD_BUF_NORMAL = 3, RPD_BUF_WORLD_POS = 4,
    RPD_BUF_DENOISED = 5,          /* latest statmc_denoise output */
    /* scalar float images */
    RPD_BUF_DEPTH = 10, RPD_BUF_UNCERTAINTY = 11, RPD_BUF_ALPHA = 12, RPD_BUF_VARIANCE_DENOISED = 13,
    RPD_BUF_VAR_TOTAL = 14, RPD_BUF_VAR_EFF = 15, RPD_BUF_VAR_RATIO = 16,
    RPD_BUF_SENSITIVITY_CONFIDENCE = 17, RPD_BUF_SENSITIVITY_GRADIENT = 18, RPD_BUF_ADAPTIVE_IMPORTANCE = 19,
    RPD_BUF_SAMPLE_FRACTION = 20,  /* sample_count / max(sample_count) */
    /* expanded sensitivity fields (SensitivityTiles::expand_values) */
    RPD_BUF_SENS_ALL = 30, RPD_BUF_SENS_PIXEL, RPD_BUF_SENS_BRDF, RPD_BUF_SENS_LENS, RPD_BUF_SENS_LIGHT,
    RPD_BUF_SENS_LIGHT_UV, RPD_BUF_SENS_LIGHT_SELECT, RPD_BUF_SENS_ENVIRONMENT, RPD_BUF_SENS_RR,
    /* int32 per pixel */
    RPD_BUF_SAMPLE_COUNT = 50, RPD_BUF_HIT_COUNT = 51, RPD_BUF_PASS_COUNTS = 52,
    /* raw state, readable and writable */
    RPD_BUF_STATS = 60,            /* 12 floats/pixel: lum mean, lum M2, n (int bits), color mean[3], sqrt mean[3], sqrt M2[3] */
    RPD_BUF_SUMS = 61,             /* 13 floats/pixel: radiance[3], albedo[3], normal[3], world_pos[3], depth sums */
    RPD_BUF_GRID = 62,             /* 440 doubles/node (RPF grid accumulators) */
    RPD_BUF_TILES = 63,            /* 18 planes of node floats (SensitivityTiles fields) */
    RPD_BUF_FEATURES = 64,         /* 7 floats/pixel: unit normal[3], depth, albedo[3] */
    RPD_BUF_RECORDS = 65,          /* last pass: 32 floats per sample (read only) */
    /* debug_statmc_outputs (main.cpp:991-1157, 1283-1321), kept only for jobs
       created after rpd_ctx_set_debug_outputs(ctx, 1): per pixel of the last
       StatMC filter the accepted neighbours, weight sum and effective
       neighbours (0 without a primary hit), and the clamped variance source of
       the last adaptive allocation. Floats per pixel. */
    RPD_BUF_DEBUG_NEIGHBORS = 70, RPD_BUF_DEBUG_WEIGHT_SUM = 71, RPD_BUF_DEBUG_NEFF = 72,
    RPD_BUF_DEBUG_VARIANCE_SOURCE = 73
};
/* Size in bytes of a buffer for this job (RECORDS: of the last pass). */
int rpd_job_buffer_size(rpd_job* job, int buffer_id, size_t* bytes);
int rpd_job_read(rpd_job* job, int buffer_id, void* dst, size_t bytes);
int rpd_job_write(rpd_job* job, int buffer_id, const void* src, size_t bytes);
/* Grid geometry: nodes_x, nodes_y, step, support (core/stats.h:312-318). */
int rpd_job_grid_dims(rpd_job* job, int dims[4]);

/* Stage timings in ms, measured with HIP events on the job's stream:
   [0] base render, [1] adaptive renders, [2] sensitivity, [3] StatMC,
   [4] variance smoothing, [5] allocation, [6] whole job,
   [7] trace kernel only (all passes), [8] number of trace launches,
   [9] StatMC window kernel only, [10] grid gather kernels only. */
int rpd_job_stage_ms(rpd_job* job, double* ms, int n);

/* ---- denoising caller-supplied or synthetic sample records ----------------
   The reference folds each traced sample into the pixel statistics, features
   and RPF grid inside its sample loop (main.cpp:707-717, 329-376) and then
   denoises (recompute_sensitivity, main.cpp:741; statmc_denoise, 957). These
   entry points run the same fold and denoise on per-sample records that come
   from the caller (or from the device-side synthetic generator) instead of
   from the path tracer. */
enum rpd_sample_flags {
    RPD_SAMPLE_FINITE = 1u << 0,        /* set by the engine from the radiance */
    RPD_SAMPLE_HIT = 1u << 1,           /* primary hit: normal, depth, albedo, position valid */
    RPD_SAMPLE_PIXEL = 1u << 2,         /* rpf::Sample validity bits (core/stats.h:331-399) */
    RPD_SAMPLE_LENS = 1u << 3,
    RPD_SAMPLE_LIGHT_SELECT = 1u << 4,
    RPD_SAMPLE_BRDF = 1u << 5,
    RPD_SAMPLE_LIGHT = 1u << 6,
    RPD_SAMPLE_ENV = 1u << 7,
    RPD_SAMPLE_RR = 1u << 8,
    RPD_SAMPLE_RR_SURVIVED = 1u << 9
};
/* One path sample: radiance, primary-hit features (main.cpp:572-576) and the
   first-occurrence random parameters of rpf::Sample. */
typedef struct rpd_sample_record {
    float radiance[3];
    uint32_t flags;       /* rpd_sample_flags */
    float normal[3];
    float depth;          /* primary hit distance */
    float albedo[3];
    float position[3];
    float pixel_u[2];
    float lens_u[2];
    float brdf_u[2];
    float light_u[2];
    float env_u[2];
    int32_t light_index;  /* light selected first; bins clamp it to [0, 15] */
} rpd_sample_record;

/* Synthetic records: square regions of region_size pixels, each a plane with
   its own unit normal, depth uniform in [1, 20] and albedo uniform in
   [0.1, 0.9]; lognormal radiance exp(log_mean + log_sigma * z) tinted by the
   albedo, a firefly_fraction of samples scaled by firefly_scale; ten random
   dimensions uniform in [0, 1], three of them (lens x, brdf x, light x) tied to
   z through a Gaussian copula with the given correlation. */
typedef struct rpd_synth_config {
    uint32_t seed;
    int samples_per_pixel;
    int region_size;
    float log_mean, log_sigma, firefly_fraction, firefly_scale, correlation;
} rpd_synth_config;
void rpd_synth_config_defaults(rpd_synth_config* c); /* mu -1, sigma 1.5, 2% fireflies x100, rho 0.6, 32 px */

/* A job for denoising records only: no scene, no camera, no sampling tape;
   room for max_samples_per_pixel samples of every pixel per pass. */
int rpd_job_create_denoiser(rpd_context* ctx, const rpd_render_config* cfg, int max_samples_per_pixel,
                            rpd_job** out);
/* Loads one pass of records: counts[p] samples of pixel p (row-major), the
   records in pixel-major order, n = sum of counts. */
int rpd_job_load_records(rpd_job* job, const int* counts, const rpd_sample_record* records, size_t n);
/* Writes one pass of synthetic records on the device (uniform samples per pixel). */
int rpd_job_synthesize_records(rpd_job* job, const rpd_synth_config* sc);
/* Folds the loaded pass into the pixel statistics, features and RPF grid,
   then recompute_sensitivity and statmc_denoise; RPD_BUF_DENOISED then holds
   the result. */
int rpd_job_denoise_records(rpd_job* job);
/* The same as rpd_job_denoise_records for n jobs at once: each stage is one
   launch over all images (persistent kernels for the grid gather and the
   StatMC window). Results are identical to denoising each job alone.
   gpu_ms (optional) receives the device time of the whole batch. */
int rpd_denoise_batch(rpd_context* ctx, rpd_job* const* jobs, int n, double* gpu_ms);

/* ---- procedural benchmark scene (no counterpart in the reference) ---------
   An icosphere subdivided subdivision_level times (20 * 4^level flat
   triangles; level 8 gives 1,310,720) with its vertices displaced along the
   radius by noise_octaves octaves of value noise, n_spheres spheres with GGX
   microfacet materials (roughness uniform in [0.05, 0.8], albedo uniform in
   [0.1, 0.9]) in a shell around it, four rectangular area lights facing it and
   a constant environment. The same configuration always gives the same bits. */
typedef struct rpd_procedural_config {
    uint32_t seed;
    int subdivision_level;   /* 0..9 */
    int n_spheres;
    float displacement;      /* radial displacement amplitude (unit sphere) */
    float noise_frequency;
    int noise_octaves;
    float environment[3];    /* constant environment radiance */
    float light_emission;
} rpd_procedural_config;
void rpd_procedural_config_defaults(rpd_procedural_config* c); /* seed 7, level 8, 256 spheres */
/* The scene as arrays for rpd_scene_create. Call with null arrays to get the
   counts; then with arrays of at least those sizes. */
int rpd_procedural_scene_arrays(const rpd_procedural_config* c, rpd_triangle* tris, int* n_tris,
                                rpd_sphere* spheres, int* n_spheres, rpd_material* mats, int* n_mats,
                                rpd_light* lights, int* n_lights, float env[3]);
/* ctx == NULL (here and in rpd_scene_create) keeps the scene on the host, for
   rpd_scene_query only. */
int rpd_scene_procedural(rpd_context* ctx, const rpd_procedural_config* c, rpd_scene** out);

/* Ray queries against a scene on the host, with the same intersection code
   the kernels run: use_bvh != 0 walks the scene's BVH (scenes that have one),
   0 loops over every primitive. occlusion[i] != 0 asks whether anything lies
   before t_max[i] (t_max already reduced by 5e-4 as in any_hit); otherwise the
   closest hit below t_max[i]. Closest-hit and any-hit semantics are those of
   core/ray_tracer.h:233-344. For checking the BVH against the exhaustive loop. */
int rpd_scene_query(const rpd_scene* s, int n, const float* origins, const float* dirs, const float* t_max,
                    const int* occlusion, int use_bvh, int* hit, float* t, int* index, int* is_sphere);
/* Nodes and 4-wide depth of the scene's BVH (0, 0 for scenes looped over). */
int rpd_scene_bvh_info(const rpd_scene* s, int* nodes, int* depth);
/* Which builder made the scene's BVH: RPD_BVH_HOST_SAH or RPD_BVH_DEVICE_LBVH. */
int rpd_scene_bvh_builder(const rpd_scene* scene, int* builder);

/* StatMC window tile shapes (no counterpart in the reference, which filters
   pixel by pixel on the CPU, main.cpp:998-1127). The window kernel runs one
   thread per pixel over tiles of one of the shapes rpd_statmc_tile_shapes
   lists ((w, h) pairs; *count receives how many there are). By default a
   context tunes: the first StatMC launch of each (window radius, image-size
   class) times every shape twice on its own images (in order, then in
   reverse) and keeps the one with the fastest run.
   rpd_ctx_set_statmc_tile fixes one shape instead; 0, 0 returns to tuning.
   Every shape gives the same bits. rpd_ctx_statmc_tile reports the shape used
   for a radius and pixel count (0, 0: not tuned yet). */
int rpd_statmc_tile_shapes(int* wh, int n, int* count);
int rpd_ctx_set_statmc_tile(rpd_context* ctx, int tile_w, int tile_h);
int rpd_ctx_statmc_tile(rpd_context* ctx, int radius, int64_t pixels, int* tile_w, int* tile_h);

/* Jobs created after this call with on != 0 keep the debug buffers of the
   reference's debug_statmc_outputs (RPD_BUF_DEBUG_*); 0 stops it. */
int rpd_ctx_set_debug_outputs(rpd_context* ctx, int on);

/* Capture each job's adaptive-pass sequence (variance smoothing, allocation,
   sampling pass) as a hipGraph on its first pass and replay it for every
   pass; results are identical to launching the kernels one by one. */
int rpd_ctx_set_graph_passes(rpd_context* ctx, int on);
/* Extension (no counterpart in the reference, whose sample loop is
   main.cpp:649-737 and 1451-1538 around shade_mis, main.cpp:394-533): passes
   over scenes traced without a BVH (the flat triangle loop) run as a wavefront
   trace of ceil(rounds_per_sample * samples per pixel) rounds, each a shading
   launch over path slots held in device memory (one bounce or camera ray per
   slot, its queries listed) and an intersection launch over the listed
   queries; the path-regeneration kernel then finishes the slots still
   running. Records are the same as with the path-regeneration kernel alone,
   which 0 (the default) selects. rounds_per_sample must lie in [0, 64]. It
   applies to jobs created after the call: their slabs hold the path slots
   (51 state words per pixel, up to 4M slots, plus 3 query words and the round
   counters), which rpd_job_bytes does not count. */
int rpd_ctx_set_trace_wavefront(rpd_context* ctx, float rounds_per_sample);
/* Extension (no counterpart in the reference, whose passes splat into the RPF
   grid inside the sample loop, main.cpp:329-388 and 1451-1538): with on != 0
   (the default), jobs created after the call over scenes traced without a BVH
   (the flat triangle loop) whose adaptive passes gather into the grid (StatMC
   on, adaptive passes, rpf_shrinkage_scale > 0) run each
   pass's grid gather, the last excepted, on a second stream beside the next
   pass's trace, the passes alternating between two plan and record sets; the
   second set holds (adaptive_spp + 1) samples per pixel. Outputs are the same
   bits as with every gather in line, which 0 selects (and graph passes
   always use). */
int rpd_ctx_set_gather_overlap(rpd_context* ctx, int on);
/* StatMC window (main.cpp:1032-1097) in two passes, on (the default) or off
   (one pass). A pair's weight is the same bits with its two pixels exchanged,
   so the pair pass computes each pair of the window once into a table held in
   the job's tape buffer (free once the passes are done; radius up to 13) and
   the sum pass adds each pixel's full window in the reference's order. Same
   bits as one pass. Batches take it with the tables in a context workspace
   of at most 16 GiB (allocated on first use, freed with the context);
   denoise-only jobs denoised alone and radii above 13 take one pass. */
int rpd_ctx_set_statmc_pairs(rpd_context* ctx, int on);
/* Device bytes a job of this configuration needs (its slab size; admission by
   HBM footprint), the second plan and record set of the gather overlap
   included wherever its configuration allows it (an upper bound for scenes
   traced with a BVH). A job that takes a larger kept slab holds that slab's
   size. */
int rpd_job_bytes(const rpd_render_config* cfg, uint64_t* bytes);
/* Bytes held by the context's jobs, its budget (0 = none) and the device's free and total memory. */
int rpd_ctx_memory(rpd_context* ctx, uint64_t* in_use, uint64_t* budget, uint64_t* device_free,
                   uint64_t* device_total);

/* ---- serving ---------------------------------------------------------------
   A request loop over render+denoise jobs (no counterpart in the reference,
   which renders one image per process). Requests are queued as they arrive
   and admitted, first in first out with a short look-ahead, whenever a slot is
   free and the job's device footprint fits the memory budget; every admitted
   job runs its whole pipeline on its own stream, so the work of all admitted
   jobs overlaps on the device, and keeps its state resident across its
   adaptive passes. A job's result does not depend on what else runs. */
typedef struct rpd_server rpd_server;
typedef struct rpd_server_config {
    int max_in_flight;       /* jobs admitted at once (>= 1) */
    uint64_t memory_budget;  /* bytes of job state held at once, counting finished jobs until they are
                                released (0 = free device memory at creation) */
    int graph_passes;        /* replay the adaptive passes from a captured hipGraph */
    int cobatch_tail;        /* run the denoise tails of the jobs whose passes finished together as one
                                batch (rpd_render_finish_batch) instead of each on its own stream */
} rpd_server_config;
typedef struct rpd_server_stats {
    uint64_t submitted, completed, failed;
    int max_in_flight_seen;
    uint64_t peak_bytes;
    double latency_p50_ms, latency_p99_ms, latency_mean_ms, latency_max_ms;  /* submit -> done */
    double queue_p50_ms, queue_p99_ms;                                        /* submit -> admitted */
    uint64_t tail_batches, tail_jobs;  /* cobatch_tail: denoise-tail batches run, and the jobs in them */
    uint64_t over_budget;  /* requests admitted beyond the budget: nothing ran, nothing queued fitted
                              (finished jobs not yet released held the memory), and a caller waited
                              for the request, or the server was stopping */
} rpd_server_stats;
/* Also sets the context's slab cache (rpd_ctx_set_buffer_cache) to what the
   budget leaves of the device less 16 GiB, and at least 32 GiB, and its held
   limit (rpd_ctx_set_held_limit) to the budget. */
int rpd_server_create(rpd_context* ctx, const rpd_server_config* cfg, rpd_server** out);
/* Queues one job; returns at once with the request id. */
int rpd_server_submit(rpd_server* s, rpd_scene* scene, const rpd_render_config* cfg, const rpd_camera* cam,
                      uint32_t seed, uint64_t* request_id);
/* Blocks until the request is done; hands over its job (release it with
   rpd_job_release: its device memory counts against the budget until then)
   and its latency from submit to completion. When nothing runs and the memory
   is held by finished jobs, a request that a caller waits for is admitted
   over the budget, so a caller that keeps the jobs it claimed cannot stall.
   A failure on the scheduler thread is reported here, with its message. */
int rpd_server_wait(rpd_server* s, uint64_t request_id, rpd_job** job, double* latency_ms);
int rpd_server_stats_get(rpd_server* s, rpd_server_stats* out);
/* Finishes queued and running requests, releases jobs nobody claimed. */
int rpd_server_destroy(rpd_server* s);

/* ---- image-quality metrics (shared/tools/eval_tools.cpp) -------------------
   The reference computes these on the CPU from .hdr files: hdr_metrics
   (eval_tools.cpp:215-285: MAE, RMSE, PSNR against the reference's peak,
   global per-channel SSIM) and compute_extended_metrics (120-213: NRMSE,
   LogSSIM11 over 11x11 windows of log1p luminance, LogGradNRMSE). relMSE,
   mean over pixels and channels of (x - r)^2 / (r^2 + 1e-2), is added for
   equal-quality comparisons. Images are float RGB, interleaved, row-major;
   all arithmetic is double. */
typedef struct rpd_image_metrics {
    double mae, rmse, psnr, peak;
    double ssim[3], ssim_mean;
    double nrmse, log_ssim11, log_gradient_nrmse;
    double relmse;
} rpd_image_metrics;
/* `test` against `ref`, both width x height; on_device: the pointers are
   device memory of the context's device (else host memory, copied in). */
int rpd_image_metrics_compute(rpd_context* ctx, const float* ref, const float* test, int width, int height,
                              int on_device, rpd_image_metrics* out);
/* No counterpart in the reference. rpd_image_metrics_compute for images in
   device memory, for callers that cannot allocate device memory (code under
   stream capture, tight serving loops, harnesses that forbid allocation): it
   allocates nothing. Its partial sums live in a fixed device array of the
   library (5.2 MB, held by every device the library runs on until the process
   ends), so images are limited to 65536 16x16 tiles (4096 x 4096 pixels);
   larger ones are refused with RPD_ERR_INVALID_ARG. Calls on one context are
   serialised; calls through different contexts on the same device share the
   array and must not overlap. Results equal rpd_image_metrics_compute's. */
int rpd_image_metrics_compute_noalloc(rpd_context* ctx, const float* ref, const float* test, int width, int height,
                                      rpd_image_metrics* out);
/* A job's RPD_BUF_RADIANCE / RPD_BUF_RAW_RADIANCE (finalized jobs) or
   RPD_BUF_DENOISED buffer, read in place on the device, against `ref`. */
int rpd_job_image_metrics(rpd_job* job, int buffer_id, const float* ref, int ref_on_device, rpd_image_metrics* out);

/* ---- comparison denoisers (denoising/joint_bilateral.cpp, atrous_wavelet.cpp)
   The reference's two baseline filters, which it runs as separate tools on
   .hdr files. Images are float RGB, interleaved, row-major, as those tools
   read them; outputs must not alias inputs. on_device: every pointer is device
   memory of the context's device (else host memory). Float arithmetic in the
   reference's order. */
typedef struct rpd_bilateral_config {
    float sigma_c, sigma_n, sigma_p, sigma_a;  /* colour, normal angle (radians), pixel distance, albedo */
} rpd_bilateral_config;
void rpd_bilateral_config_defaults(rpd_bilateral_config* c);  /* all 1 (joint_bilateral.cpp:172-181) */
/* joint_bilateral.cpp:15-128: Gaussian, colour bilateral and joint bilateral
   images from one sweep over windows of radius max(1, ceil(3 sigma_p)) with
   clamped coordinates; `normal` is encoded as (n + 1) / 2 and decoded as
   there. Any of the three outputs may be NULL; sigma_p at most 32. */
int rpd_joint_bilateral(rpd_context* ctx, const float* raw, const float* normal, const float* albedo, int width,
                        int height, const rpd_bilateral_config* cfg, float* gaussian, float* bilateral, float* joint,
                        int on_device);
typedef struct rpd_atrous_config {
    float sigma_c, sigma_n, sigma_a;  /* luminance, normal angle (radians), albedo */
    int iterations;                   /* passes at step 1, 2, 4, ... (1..30) */
} rpd_atrous_config;
void rpd_atrous_config_defaults(rpd_atrous_config* c);  /* 0.5, 0.5, 0.1, 5 iterations (atrous_wavelet.cpp:220-230) */
/* atrous_wavelet.cpp:57-183: edge-avoiding a-trous passes of the 5x5
   B3-spline kernel; normals are normalised as stored. */
int rpd_atrous(rpd_context* ctx, const float* raw, const float* normal, const float* albedo, int width, int height,
               const rpd_atrous_config* cfg, float* out, int on_device);

/* sizeof of the ABI structs, for bindings to check their mirrors: render
   config, camera, triangle, sphere, material, light, sample record, synth
   config, procedural config, server config, server stats, image metrics,
   bilateral config, a-trous config (in that order). */
int rpd_abi_struct_sizes(int* out, int n);

/* Work counters of the job so far: [0] ray queries traced (camera, extension
   and shadow rays), [1] path samples traced, [2] scene triangles, [3] scene
   spheres. Used to express the trace kernel's work in intersection tests.
   No counterpart in the reference (it reports stage times only). */
int rpd_job_counters(rpd_job* job, uint64_t* out, int n);

#ifdef __cplusplus
}
#endif
#endif /* RPD_ENGINE_H */
