// Sampling pass: MT19937 tapes, the path-regeneration trace kernel, the pass
// plan (counts -> slot offsets + active pixel list) and the per-pixel fold.
//
// Reference: path_tracer/main.cpp:394-585 (shade_mis, mis_path_trace),
// 590-737 and 1451-1538 (per-pixel sample loops), core/sampler.h,
// core/stats.h:25-41 (Welford fold).
#include "kernels.h"

#include <algorithm>

namespace rpd {

// ---------------------------------------------------------------------------
// Tape generation: one thread per active pixel seeds std::mt19937 with
// pixel_sample_seed(idx, n) and performs the first twist. The seeding words
// x[k] come from the recurrence x[k] = 1812433253 * (x[k-1] ^ x[k-1] >> 30) + k,
// walked by two cursors (at k+1 and at k+397) so the seeded state is never
// stored; word k >= 227 of the twist reads back the tape word k - 227 it wrote
// earlier. Words are written four at a time (one uint4 per lane, so a wave
// writes one contiguous KiB) and read back the same way: the chunk holding
// words 4c-227 .. 4c-224 is read one chunk ahead of its use, so the read-back
// latency overlaps the next four words' work. Only the pixel's prefix is
// written (tape_prefix_words of its sample count): an adaptive pass's pixel of
// a few samples draws a small part of the 624 words, and the trace extends the
// rare tape drawn past its prefix. A whole tape also gets the next state's
// first MT_S2 words, so that a pixel drawing a little more than 624 numbers in
// a pass (about 4% of them in a 16-sample pass) needs no twist of its own.
__device__ __forceinline__ uint32_t u4_at(const uint4& v, int i) {
    return i == 0 ? v.x : (i == 1 ? v.y : (i == 2 ? v.z : v.w));
}

__global__ void __launch_bounds__(256) k_tape(TraceParams p) {
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= p.plan.totals[1]) return;
    const int pix = p.plan.active[a];
    const int n_before = p.seed_n ? p.seed_n[pix] : 0;
    const uint32_t s = pixel_sample_seed(p.base_seed, pix, n_before);
    uint32_t* t = p.tape + tape_base(a);
    const int n_chunks = tape_prefix_words(p.plan.counts[pix], p.tape_per_sample, p.tape_margin) / 4;
    const bool whole = n_chunks == MT_N / 4;
    uint32_t lo = s;               // x[k]
    uint32_t far = s;              // x[k + 397] while k + 397 < 624
    for (uint32_t k = 1; k <= (uint32_t)MT_M; ++k) far = mt_seed_step(far, k);
    constexpr int kBack = MT_N - MT_M;  // 227: word k >= kBack mixes in tape word k - kBack
    constexpr int kLastBackChunk = (MT_N - 2 - kBack) / 4;  // 98: chunk of the last such word (622 - 227)
    // Window of read-back chunks: at the start of iteration c, m57 = chunk c-57
    // and m56 = chunk c-56 (word k = 4c+i reads word 4(c-57)+i+1); chunk c-54 is
    // requested at the end of iteration c, after the iteration that wrote it.
    uint4 m57 = make_uint4(0, 0, 0, 0), m56 = make_uint4(0, 0, 0, 0), m55 = make_uint4(0, 0, 0, 0);
    // The next state's first MT_S2 words, for a whole tape: its word k mixes
    // words k, k+1 and k+397 of this state, so chunk m of it is formed when
    // this state's chunk m+100 is, from chunks m and m+1 (read back) and
    // m+99 and m+100 (in registers).
    constexpr int kS2From = MT_M / 4 + 1;                        // 100: chunk of word 397 + 3
    uint4 prev = make_uint4(0, 0, 0, 0), s2a = make_uint4(0, 0, 0, 0);
    for (int c = 0; c < n_chunks; ++c) {
        uint32_t w4[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int k = 4 * c + i;
            if (k < kBack) {                             // far word from the seeded state
                const uint32_t x1 = mt_seed_step(lo, (uint32_t)(k + 1));
                w4[i] = mt_mix(lo, x1, far);
                lo = x1;
                if (k + 1 < kBack) far = mt_seed_step(far, (uint32_t)(k + 1 + MT_M));
            } else if (k < MT_N - 1) {                   // far word: tape word k - 227
                const uint32_t back = i < 3 ? u4_at(m57, i + 1) : m56.x;
                const uint32_t x1 = mt_seed_step(lo, (uint32_t)(k + 1));
                w4[i] = mt_mix(lo, x1, back);
                lo = x1;
            } else {                                     // last word wraps to the new word 0
                w4[i] = mt_mix(lo, t[tape_word(0)], t[tape_word(MT_M - 1)]);
            }
        }
        *reinterpret_cast<uint4*>(t + tape_word(4 * c)) = make_uint4(w4[0], w4[1], w4[2], w4[3]);
        if (whole && c >= kS2From && c < kS2From + MT_S2 / 4) {
            const int m = c - kS2From;
            if (m == 0) s2a = *reinterpret_cast<const uint4*>(t + tape_word(0));
            const uint4 s2b = *reinterpret_cast<const uint4*>(t + tape_word(4 * m + 4));
            const uint32_t x2 = mt_mix(s2a.x, s2a.y, prev.y), y2 = mt_mix(s2a.y, s2a.z, prev.z);
            const uint32_t z2 = mt_mix(s2a.z, s2a.w, prev.w), w2 = mt_mix(s2a.w, s2b.x, w4[0]);
            *reinterpret_cast<uint4*>(t + tape_word(MT_S2_AT + 4 * m)) = make_uint4(x2, y2, z2, w2);
            s2a = s2b;
        }
        prev = make_uint4(w4[0], w4[1], w4[2], w4[3]);
        m57 = m56;
        m56 = m55;
        const int ahead = c - 54;
        if (ahead >= 0 && ahead <= kLastBackChunk)
            m55 = *reinterpret_cast<const uint4*>(t + tape_word(4 * ahead));
    }
    if (!whole) {   // where the trace goes on if it draws past the prefix
        t[tape_word(MT_TAPE_CURSOR)] = lo;
        t[tape_word(MT_TAPE_CURSOR + 1)] = far;
    }
}

void launch_tape(const TraceParams& p, hipStream_t st) {
    const int blocks = (p.plan.n_pixels + 255) / 256;
    hipLaunchKernelGGL(k_tape, dim3(blocks), dim3(256), 0, st, p);
}

// ---------------------------------------------------------------------------
// Trace kernel. One lane owns one pixel at a time and draws that pixel's
// samples in order from its tape, so every random number lands where the
// reference draws it. Lanes regenerate: a lane whose sample ended starts the
// pixel's next sample, and a lane whose pixel ended takes the next pixel from
// a wave-level chunk of the active list.
//
// Every loop iteration resolves the lane's queries of the last iteration,
// then shades one bounce in full (main.cpp:394-533): light selection and light
// sample, environment sample, BRDF sample, in the reference's draw order, and
// issues that bounce's rays together: the light's shadow ray, the environment
// shadow ray and the BRDF (extension) ray. All random draws of a bounce are
// independent of the shadow-ray outcomes, and the Russian-roulette draw
// follows the extension ray's result, so the draw order is the reference's.
// Because every lane shades at most one bounce per iteration, the lanes of a
// wave run the same shading code together instead of splitting between a
// light step and a BRDF step.
//
// Radiance is accumulated forward (throughput times each bounce's direct
// term); the reference's recursion sums the same terms innermost-first.
enum : int { ST_NEED = 0, ST_CAMERA = 1, ST_SHADE = 2, ST_WAIT = 3, ST_DONE = 5, ST_DEAD = 6 };
// queries of one iteration (bit mask): the closest-hit ray (camera or BRDF
// extension), the light's shadow ray, the environment's shadow ray
enum : int { Q_EXT = 1, Q_SHADOW_L = 2, Q_SHADOW_E = 4, Q_PRIMARY = 8 };

constexpr int TRACE_BLOCK = 256;
constexpr int TRACE_CHUNK = 64;

__device__ __forceinline__ uint64_t ballot64(bool b) { return __ballot(b); }
__device__ __forceinline__ bool is_zero3(v3 a) { return a.x == 0.0f && a.y == 0.0f && a.z == 0.0f; }

// Issue priority of a wave on its SIMD, from the most samples one of its lanes
// still has to trace. A pixel's samples run one after another on one lane, so
// once the active list is used up the pass ends with the wave that holds the
// longest remaining chain; with a higher priority its instructions issue ahead
// of those of waves with shorter chains, which fill the cycles it leaves. The
// tests are ballots (wave-uniform), so exactly one s_setprio runs.
__device__ __forceinline__ void wave_priority(int left) {
    if (__ballot(left >= 16)) __builtin_amdgcn_s_setprio(3);
    else if (__ballot(left >= 8)) __builtin_amdgcn_s_setprio(2);
    else if (__ballot(left >= 4)) __builtin_amdgcn_s_setprio(1);
    else __builtin_amdgcn_s_setprio(0);
}

// Scenes of up to TRACE_LDS_TRIS triangles, TRACE_FLAT_SPHERES spheres and
// TRACE_LDS_MATERIALS materials and lights each are traced by the flat loop,
// with their triangle tables (the permuted geometry and the shading
// attributes), materials, lights and light CDF staged in LDS: at most 39.2 KB
// per 256-lane block, so four blocks fit a CU. Larger scenes go through the
// BVH and read their tables from global memory.
constexpr int TRACE_LDS_TRIS = 192;
constexpr int TRACE_FLAT_SPHERES = 32;
constexpr int TRACE_LDS_MATERIALS = 64;

// The rays a lane issues in one iteration: all start at `ro` (the camera, or
// the offset shading point), with their own directions.
struct LaneRays {
    v3 ro;
    v3 rd_ext, rd_l, rd_e;   // closest-hit ray, light shadow ray, environment shadow ray
    float t_l;               // light shadow ray range (distance - any-hit epsilon)
};

// One step of a path lane: resolves the lane's queries (h: the closest-hit
// result, occ_l / occ_e: whether the shadow rays were blocked) once none is
// pending, writes a finished sample, hands out pixels (a wave-wide operation),
// then starts a camera ray or shades one bounce, setting q and pend to the
// queries to run. Returns false once every lane of the wave is out of work.
template <bool kDepth, bool kLambert>
__device__ __forceinline__ bool path_step(const TraceParams& p, const DevScene& S, const int lane, const v3 env, const bool env_on,
                                          int& c_next, int& c_end, const int n_active, const HitRec& h,
                                          const bool occ_l, const bool occ_e, int& st, int& q, int& pend,
                                          int& pix, int& left,
                                          int64_t& slot, Draws& dr, LaneRays& R, v3& T, v3& L, v3& Ld, v3& sp,
                                          v3& sn, v3& swo, v3& cw_l, v3& cw_e, v3& cw_x, float& bpdf, int& smat,
                                          int& depth, uint32_t& flags, uint32_t& light_bits) {
    const float pdf_env = float(1.0 / (4.0 * RPD_PI));
    // -- resolve the last queries -------------------------------------------
    if (st == ST_WAIT && pend == 0) {
        if (q & Q_PRIMARY) {
            // main.cpp:540-585
            st = ST_DONE;
            if (!h.hit) {
                L = env;
            } else {
                bool emitter;
                int li, hm;
                v3 hn;
                if (h.sphere) {
                    const DevSphere& sp_ = S.spheres[h.index];
                    emitter = sp_.emitter != 0; li = sp_.light; hn = h.sn; hm = sp_.material;
                } else {
                    const TriShade t = tri_attr_unpack(S.tri_attr[h.index]);
                    emitter = t.emitter; li = t.light; hn = t.n; hm = t.material;
                }
                const v3 hp = R.ro + h.t * R.rd_ext;
                flags |= REC_HIT;
                p.rec.r1[slot] = make_float4(hn.x, hn.y, hn.z, h.t);
                const DevMaterial& hmat = S.materials[hm];
                p.rec.r2[slot] = make_float4(hmat.albedo[0], hmat.albedo[1], hmat.albedo[2], 0.0f);
                p.rec.r3[slot] = make_float4(hp.x, hp.y, hp.z, 0.0f);
                if (emitter) {
                    L = eval_light(S.lights[li], -R.rd_ext, h.t);
                } else {
                    sp = hp; sn = hn; swo = -normalized3(R.rd_ext); smat = hm;
                    depth = 1;
                    st = ST_SHADE;
                }
            }
        } else {
            // the shadow rays of the bounce (main.cpp:436-448, 456-461)
            if ((q & Q_SHADOW_L) && !occ_l) Ld = Ld + cw_l;
            if ((q & Q_SHADOW_E) && !occ_e) Ld = Ld + cw_e;
            st = ST_DONE;
            if (!(q & Q_EXT)) {
                L = L + vmul(T, Ld);      // no BRDF ray: the bounce ends the path
            } else {
                // the BRDF ray: main.cpp:480-529
                const DevMaterial& m = S.materials[smat];
                const bool delta_brdf = !kLambert && m.kind == MAT_MIRROR;
                bool emitter = false;
                int li = -1, hm = 0;
                v3 hn = mk3(0, 0, 0);
                if (h.hit) {
                    if (h.sphere) {
                        const DevSphere& sp_ = S.spheres[h.index];
                        emitter = sp_.emitter != 0; li = sp_.light; hn = h.sn; hm = sp_.material;
                    } else {
                        const TriShade t = tri_attr_unpack(S.tri_attr[h.index]);
                        emitter = t.emitter; li = t.light; hn = t.n; hm = t.material;
                    }
                }
                if (h.hit && emitter) {
                    const DevLight& E = S.lights[li];
                    float wb = 1.0f;
                    if (!delta_brdf && !light_is_delta(E)) {
                        const float plb = light_pdf(S, li, light_select_pdf(S, li), R.rd_ext, h.t);
                        if (plb > RPD_EPS_SMALL) wb = bpdf / (bpdf + plb);
                    }
                    Ld = Ld + wb * vmul(eval_light(E, -R.rd_ext, h.t), cw_x);
                } else if (!h.hit) {
                    const float wb = delta_brdf ? 1.0f : bpdf / (bpdf + pdf_env);
                    Ld = Ld + wb * vmul(env, cw_x);
                }
                L = L + vmul(T, Ld);
                if (h.hit && !emitter) {
                    const float ksi = dr.next();
                    if (p.capture && !(flags & REC_RR)) {
                        flags |= REC_RR;
                        if (ksi < 0.8f) flags |= REC_RR_SURVIVED;
                    }
                    if (ksi < 0.8f) {
                        T = vmul(T, cw_x) / 0.8f;
                        sp = R.ro + h.t * R.rd_ext;
                        sn = hn;
                        swo = -R.rd_ext;
                        smat = hm;
                        ++depth;
                        st = ST_SHADE;
                    }
                }
            }
        }
        q = 0;
    }

    // -- finish samples ------------------------------------------------
    if (st == ST_DONE) {
        if (finite3(L)) flags |= REC_FINITE;
        flags |= (light_bits & 255u) << REC_LIGHT_SHIFT;
        p.rec.r0[slot] = make_float4(L.x, L.y, L.z, __uint_as_float(flags));
        ++slot;
        --left;
        st = left > 0 ? ST_CAMERA : ST_NEED;
    }
    // -- hand out pixels ------------------------------------------------
    uint64_t need = ballot64(st == ST_NEED);
    while (need) {
        if (c_next >= c_end) {
            int base = 0;
            if (lane == 0) base = atomicAdd(p.work_counter, TRACE_CHUNK);
            base = __shfl(base, 0);
            c_next = base;
            c_end = min(base + TRACE_CHUNK, n_active);
            if (c_next >= c_end) break;
        }
        const int avail = c_end - c_next;
        const int rank = __popcll(need & ((1ull << lane) - 1ull));
        const bool mine = (st == ST_NEED) && rank < avail;
        if (mine) {
            const int a = c_next + rank;
            pix = p.plan.active[a];
            left = p.plan.counts[pix];
            slot = p.plan.offsets[pix];
            dr.reset(p.tape + tape_base(a), tape_prefix_words(left, p.tape_per_sample, p.tape_margin));
            st = ST_CAMERA;
        }
        const int taken = min(__popcll(need), avail);
        c_next += taken;
        need = ballot64(st == ST_NEED);
    }
    if (st == ST_NEED) st = ST_DEAD;
    if (ballot64(st != ST_DEAD) == 0) return false;
    wave_priority(left);

    // -- camera rays: main.cpp:691-705 ----------------------------------
    if (st == ST_CAMERA) {
        const int x = pix % p.width, y = pix / p.width;
        const v2 c_pix = dr.next2();
        const float u = (float(x) + c_pix.x) / float(p.width);
        const float v = (float(y) + c_pix.y) / float(p.height);
        const v2 c_lens = dr.next2();
        v2 lens; lens.x = 0.0f; lens.y = 0.0f;
        if (p.cam.lens_radius > 0.0f) lens = sample_disk(c_lens);
        flags = 0;
        light_bits = 255u;
        if (p.capture) {
            flags |= REC_PIXEL;
            if (p.cam.lens_radius > 0.0f) flags |= REC_LENS;
            p.rec.r4[slot] = make_float4(c_pix.x, c_pix.y, c_lens.x, c_lens.y);
        }
        v3 o = ld3(p.cam.pos);
        if (p.cam.lens_radius > 0.0f) {
            const float lx = p.cam.lens_radius * lens.x, ly = p.cam.lens_radius * lens.y;
            o = o + (lx * ld3(p.cam.right) + ly * ld3(p.cam.up));
        }
        const v3 px = ld3(p.cam.llc) + u * ld3(p.cam.horiz) + (1.0f - v) * ld3(p.cam.vert);
        R.ro = o;
        R.rd_ext = normalized3(px - o);
        T = mk3(1, 1, 1);
        L = mk3(0, 0, 0);
        q = Q_EXT | Q_PRIMARY;
        pend = Q_EXT;
        st = ST_WAIT;
    }

    // -- one bounce: light sample (main.cpp:405-448), environment sample
    //    (450-462) and BRDF sample (464-479), in the reference's draw order ----
    if (st == ST_SHADE) {
        const DevMaterial& m = S.materials[smat];
        const int mkind = kLambert ? int(MAT_LAMBERT) : m.kind;
        const bool delta_brdf = !kLambert && m.kind == MAT_MIRROR;
        Ld = mk3(0, 0, 0);
        R.ro = offset_origin(sp, sn);
        const float sel_u = dr.next();
        int li; float sel_pdf;
        sample_light(S, sel_u, li, sel_pdf);
        if (p.capture && !(flags & REC_LSEL) && li >= 0 && li < 16) { flags |= REC_LSEL; light_bits = (uint32_t)li; }
        if (li >= 0 && (!delta_brdf || light_is_delta(S.lights[li]))) {
            const DevLight& lt = S.lights[li];
            const bool is_delta = light_is_delta(lt);
            const v2 lu = dr.next2();
            if (p.capture && !(flags & REC_LIGHT) && !is_delta) {
                flags |= REC_LIGHT;
                reinterpret_cast<float2*>(&p.rec.r5[slot])[1] = make_float2(lu.x, lu.y);
            }
            v3 wi; float dist;
            light_sample(lt, sp, lu, wi, dist);
            if (dist > RPD_EPS_SMALL || is_delta) {
                const float pl = light_pdf(S, li, sel_pdf, wi, dist);
                if (pl > RPD_EPS_SMALL) {
                    const float pw = fmaxf(RPD_EPS_SMALL, brdf_pdf(mkind, m, swo, wi, sn));
                    const float wl = is_delta ? 1.0f : pl / (pl + pw);
                    const v3 em = eval_light(lt, -wi, dist);
                    if (maxc3(em) > 0.0f) {
                        const v3 f = brdf_eval(mkind, m, swo, wi, sn);
                        const float g = delta_brdf ? 1.0f : fmaxf(0.0f, dot3(sn, wi));
                        const v3 contrib = (vmul(em, f) * g) / pl;
                        cw_l = wl * contrib;
                        const float eff = dist - RPD_EPS_ANYHIT;  // ray_tracer.h:65-75
                        if (eff <= RPD_EPS_SMALL) {
                            Ld = Ld + cw_l;
                        } else if (!is_zero3(cw_l)) {
                            R.rd_l = wi; R.t_l = eff; q |= Q_SHADOW_L;
                        }
                    }
                }
            }
        }
        const v2 eu = dr.next2();
        if (env_on && !delta_brdf) {
            const v3 ed = sample_sphere(eu);
            if (p.capture && !(flags & REC_ENV)) {
                flags |= REC_ENV;
                p.rec.r6[slot] = make_float2(eu.x, eu.y);
            }
            const float bp = brdf_pdf(mkind, m, swo, ed, sn);
            const float we = pdf_env / (pdf_env + bp);
            const v3 f = brdf_eval(mkind, m, swo, ed, sn);
            const float c = fmaxf(0.0f, dot3(sn, ed));
            cw_e = (we * vmul(env, f)) * (c / pdf_env);
            if (!is_zero3(cw_e)) { R.rd_e = ed; q |= Q_SHADOW_E; }
        }
        const v2 bu = dr.next2();
        if (p.capture && !delta_brdf && !(flags & REC_BRDF)) {
            flags |= REC_BRDF;
            reinterpret_cast<float2*>(&p.rec.r5[slot])[0] = make_float2(bu.x, bu.y);
        }
        v3 dir; float pdf;
        brdf_sample(mkind, m, swo, sn, bu, dir, pdf);
        if (pdf > RPD_EPS_SMALL && sqnorm3(dir) > 0.0f && (!kDepth || p.max_depth <= 0 || depth < p.max_depth)) {
            bpdf = pdf;
            cw_x = brdf_sample_weight(mkind, m, swo, dir, sn, pdf);
            R.rd_ext = dir;
            q |= Q_EXT;
        }
        // a bounce with no query ends here; it is written next iteration
        pend = q & (Q_EXT | Q_SHADOW_L | Q_SHADOW_E);
        st = ST_WAIT;
    }
    return true;
}

// ---------------------------------------------------------------------------
// Path slots of the wavefront trace. A slot holds one path lane's state
// between kernels: the words path_step keeps from one call to the next (the
// shading point itself is formed and used within one call), the rays of its
// queries and their results. Only a lane waiting on queries (ST_WAIT) has
// state beyond its ST word, and of that only what its issued queries need.
enum : int {
    WF_ST = 0, WF_Q, WF_PIX, WF_LEFT, WF_TAPE, WF_POS, WF_LIMIT, WF_SLOT, WF_FLAGS, WF_LBITS, WF_DEPTH, WF_SMAT,
    WF_BPDF, WF_T = 13, WF_L = 16, WF_LD = 19, WF_CWL = 22, WF_CWE = 25, WF_CWX = 28,
    WF_RO = 31, WF_RDX = 34, WF_RDL = 37, WF_TL = 40, WF_RDE = 41,
    WF_HCODE = 44, WF_HT = 45, WF_HSN = 46, WF_OCCL = 49, WF_OCCE = 50, WF_USED = 51
};
static_assert(WF_USED <= WF_WORDS, "wavefront slot words");
__device__ __forceinline__ int* wf_count(const WavefrontState& w, int iter) { return w.ctrl + iter * WF_LISTS; }
__device__ __forceinline__ int* wf_live(const WavefrontState& w, int iter) {
    return w.ctrl + (WF_MAX_ITERS + iter) * WF_LISTS;
}

struct WfRef {
    const WavefrontState& w;
    int s;
    __device__ __forceinline__ uint32_t u(int k) const { return w.state[size_t(k) * w.slots + s]; }
    __device__ __forceinline__ float f(int k) const { return __uint_as_float(u(k)); }
    __device__ __forceinline__ v3 f3(int k) const { return mk3(f(k), f(k + 1), f(k + 2)); }
    __device__ __forceinline__ void put(int k, uint32_t v) const { w.state[size_t(k) * w.slots + s] = v; }
    __device__ __forceinline__ void putf(int k, float v) const { put(k, __float_as_uint(v)); }
    __device__ __forceinline__ void put3(int k, v3 v) const { putf(k, v.x); putf(k + 1, v.y); putf(k + 2, v.z); }
};

// Loads slot s into the lane's path variables. A waiting lane's queries have
// all been answered (pend = 0), so path_step resolves them first.
template <bool kDepth>
__device__ __forceinline__ void wf_load(const TraceParams& p, const WfRef& W, HitRec& h, bool& occ_l, bool& occ_e,
                                        int& st, int& q, int& pend, int& pix, int& left, int64_t& slot, Draws& dr,
                                        LaneRays& R, v3& T, v3& L, v3& Ld, v3& cw_l, v3& cw_e, v3& cw_x,
                                        float& bpdf, int& smat, int& depth, uint32_t& flags, uint32_t& light_bits) {
    st = int(W.u(WF_ST));
    pend = 0;
    if (st != ST_WAIT) return;
    q = int(W.u(WF_Q));
    pix = int(W.u(WF_PIX));
    left = int(W.u(WF_LEFT));
    slot = int64_t(W.u(WF_SLOT));
    flags = W.u(WF_FLAGS);
    light_bits = W.u(WF_LBITS);
    if (kDepth) depth = int(W.u(WF_DEPTH));
    dr.tape = p.tape + W.u(WF_TAPE);
    dr.pos = int(W.u(WF_POS));
    dr.limit = int(W.u(WF_LIMIT));
    if (dr.pos & 3) dr.cache = *reinterpret_cast<const uint4*>(dr.tape + tape_pos_word(dr.pos & ~3));
    T = W.f3(WF_T);
    L = W.f3(WF_L);
    if (!(q & Q_PRIMARY)) {
        Ld = W.f3(WF_LD);
        smat = int(W.u(WF_SMAT));
    }
    if (q & Q_EXT) {
        R.ro = W.f3(WF_RO);
        R.rd_ext = W.f3(WF_RDX);
        const int code = int(W.u(WF_HCODE));
        h.hit = code >= 0 ? 1 : 0;
        h.sphere = code >= 0 ? (code & 1) : 0;
        h.index = code >= 0 ? (code >> 1) : -1;
        h.t = W.f(WF_HT);
        if (h.sphere) h.sn = W.f3(WF_HSN);
        if (!(q & Q_PRIMARY)) {
            cw_x = W.f3(WF_CWX);
            bpdf = W.f(WF_BPDF);
        }
    }
    if (q & Q_SHADOW_L) { cw_l = W.f3(WF_CWL); occ_l = W.u(WF_OCCL) != 0; }
    if (q & Q_SHADOW_E) { cw_e = W.f3(WF_CWE); occ_e = W.u(WF_OCCE) != 0; }
}

// Stores what wf_load reads back and the rays of the issued queries.
template <bool kDepth>
__device__ __forceinline__ void wf_store(const TraceParams& p, const WfRef& W, int st, int q, int pix, int left,
                                         int64_t slot, const Draws& dr, const LaneRays& R, v3 T, v3 L, v3 Ld,
                                         v3 cw_l, v3 cw_e, v3 cw_x, float bpdf, int smat, int depth, uint32_t flags,
                                         uint32_t light_bits) {
    W.put(WF_ST, uint32_t(st));
    if (st != ST_WAIT) return;
    W.put(WF_Q, uint32_t(q));
    W.put(WF_PIX, uint32_t(pix));
    W.put(WF_LEFT, uint32_t(left));
    W.put(WF_SLOT, uint32_t(slot));
    W.put(WF_FLAGS, flags);
    W.put(WF_LBITS, light_bits);
    if (kDepth) W.put(WF_DEPTH, uint32_t(depth));
    W.put(WF_TAPE, uint32_t(dr.tape - p.tape));
    W.put(WF_POS, uint32_t(dr.pos));
    W.put(WF_LIMIT, uint32_t(dr.limit));
    W.put3(WF_T, T);
    W.put3(WF_L, L);
    if (!(q & Q_PRIMARY)) {
        W.put3(WF_LD, Ld);
        W.put(WF_SMAT, uint32_t(smat));
    }
    if (q & Q_EXT) {
        W.put3(WF_RO, R.ro);
        W.put3(WF_RDX, R.rd_ext);
        if (!(q & Q_PRIMARY)) {
            W.put3(WF_CWX, cw_x);
            W.putf(WF_BPDF, bpdf);
        }
    } else if (q & (Q_SHADOW_L | Q_SHADOW_E)) {
        W.put3(WF_RO, R.ro);
    }
    if (q & Q_SHADOW_L) { W.put3(WF_CWL, cw_l); W.put3(WF_RDL, R.rd_l); W.putf(WF_TL, R.t_l); }
    if (q & Q_SHADOW_E) { W.put3(WF_CWE, cw_e); W.put3(WF_RDE, R.rd_e); }
}

// kBvh: the scene's BVH with a per-lane traversal stack in LDS; otherwise the
// triangle table staged in LDS and looped over. kDepth: paths stop after
// p.max_depth shading points; without it the loop carries no depth count.
// kLambert: every material of the scene is Lambertian, so the microfacet and
// mirror code is left out (fewer registers, fewer divergent branches).
// kResume: the lanes first take over the path slots a wavefront trace left
// running (lane = slot), and the hand-out cursors of their 64-slot groups.
template <bool kBvh, bool kDepth, bool kLambert, bool kResume = false>
__global__ void __launch_bounds__(TRACE_BLOCK) __attribute__((amdgpu_waves_per_eu(3))) k_trace(TraceParams p) {
    // The flat loop's scene: the tables in LDS, the rest as given
    DevScene S = p.scene;
    extern __shared__ float4 s_dyn[];
    if (!kBvh) {
        const DevScene& G = p.scene;
        const int n_perm = G.n_tris * 9;
        const int n_mat4 = G.n_materials * int(sizeof(DevMaterial) / sizeof(float4));
        const int n_lig4 = G.n_lights * int(sizeof(DevLight) / sizeof(float4));
        float4* s_attr = s_dyn + n_perm;
        float4* s_mat = s_attr + G.n_tris;
        float4* s_lig = s_mat + n_mat4;
        float* s_cdf = reinterpret_cast<float*>(s_lig + n_lig4);
        for (int k = threadIdx.x; k < n_perm; k += TRACE_BLOCK) s_dyn[k] = G.tri_perm[k];
        for (int k = threadIdx.x; k < G.n_tris; k += TRACE_BLOCK) s_attr[k] = G.tri_attr[k];
        for (int k = threadIdx.x; k < n_mat4; k += TRACE_BLOCK)
            s_mat[k] = reinterpret_cast<const float4*>(G.materials)[k];
        for (int k = threadIdx.x; k < n_lig4; k += TRACE_BLOCK) s_lig[k] = reinterpret_cast<const float4*>(G.lights)[k];
        for (int k = threadIdx.x; k < G.n_lights; k += TRACE_BLOCK) s_cdf[k] = G.light_cdf[k];
        __syncthreads();
        S.tri_attr = s_attr;
        S.materials = reinterpret_cast<const DevMaterial*>(s_mat);
        S.lights = reinterpret_cast<const DevLight*>(s_lig);
        S.light_cdf = s_cdf;
    }
    const float4* tp = s_dyn;
    int* stack = reinterpret_cast<int*>(s_dyn) + threadIdx.x;  // entry k at stack[k * TRACE_BLOCK]
    if (p.query_count && blockIdx.x == 0 && threadIdx.x == 0)
        atomicAdd(p.query_count + 1, (unsigned long long)p.plan.totals[0]);  // samples of this pass
    const int lane = threadIdx.x & 63;
    const v3 env = mk3(S.env[0], S.env[1], S.env[2]);
    const bool env_on = maxc3(env) > 0.0f;
    const float inf = __builtin_inff();

    // wave-uniform chunk cursor over the active list
    int c_next = 0, c_end = 0;
    const int n_active = p.plan.totals[1];

    int st = ST_NEED, q = 0, pend = 0;
    int pix = -1, left = 0;
    int64_t slot = 0;
    Draws dr;
    dr.tape = nullptr; dr.pos = 0; dr.limit = MT_N; dr.cache = make_uint4(0, 0, 0, 0);
    LaneRays R;
    R.ro = mk3(0, 0, 0); R.rd_ext = R.ro; R.rd_l = R.ro; R.rd_e = R.ro; R.t_l = 0.0f;
    v3 T = mk3(1, 1, 1), L = mk3(0, 0, 0), Ld = mk3(0, 0, 0);
    v3 sp = mk3(0, 0, 0), sn = mk3(0, 0, 0), swo = mk3(0, 0, 0);  // shading point
    v3 cw_l = mk3(0, 0, 0), cw_e = mk3(0, 0, 0), cw_x = mk3(0, 0, 0);  // shadow contributions, BRDF weight
    float bpdf = 0.0f;
    int smat = 0, depth = 0;
    uint32_t flags = 0, light_bits = 255u;
    HitRec h; h.hit = 0; h.sphere = 0; h.index = -1; h.t = 0.0f; h.sn = mk3(0, 0, 0);
    bool occ_l = false, occ_e = false;
    uint32_t n_queries = 0;
    if (kResume) {
        const int s = blockIdx.x * TRACE_BLOCK + threadIdx.x;
        if ((s >> 6) < wf_groups(p.wf.slots)) {
            c_next = p.wf.cursor[2 * (s >> 6)];
            c_end = p.wf.cursor[2 * (s >> 6) + 1];
        }
        if (s < p.wf.slots) {
            const WfRef W{p.wf, s};
            wf_load<kDepth>(p, W, h, occ_l, occ_e, st, q, pend, pix, left, slot, dr, R, T, L, Ld, cw_l, cw_e, cw_x,
                            bpdf, smat, depth, flags, light_bits);
        }
    }

    while (true) {
        if (!path_step<kDepth, kLambert>(p, S, lane, env, env_on, c_next, c_end, n_active, h, occ_l, occ_e, st, q,
                                         pend, pix, left, slot, dr, R, T, L, Ld, sp, sn, swo, cw_l, cw_e, cw_x,
                                         bpdf, smat, depth, flags, light_bits)) {
            if (p.query_count) {
                unsigned long long tot = n_queries;
                for (int o = 32; o > 0; o >>= 1) tot += __shfl_xor(tot, o);
                if (lane == 0) atomicAdd(p.query_count, tot);
            }
            break;
        }
        // -- the lane's queries ------------------------------------------------
        // BVH scenes trace one pending query per lane per iteration (the
        // closest-hit ray first), so every lane has one walk in the divergent
        // traversal loop; the flat loop over the LDS triangle table costs the
        // same for every ray, and runs all of a lane's queries at once.
        if (kBvh) {
            const int cur = pend & -pend;
            if (cur) {
                const v3 d = cur == Q_EXT ? R.rd_ext : (cur == Q_SHADOW_L ? R.rd_l : R.rd_e);
                const HitRec r = bvh_query(p.bvh, R.ro, d, cur == Q_SHADOW_L ? R.t_l : inf, cur != Q_EXT, stack,
                                           TRACE_BLOCK);
                if (cur == Q_EXT) h = r;
                else if (cur == Q_SHADOW_L) occ_l = r.hit != 0;
                else occ_e = r.hit != 0;
                pend &= ~cur;
                ++n_queries;
            }
        } else {
            if (pend & Q_EXT) {
                h = trace_query(S, tp, R.ro, R.rd_ext, inf, false);
                ++n_queries;
            }
            if (pend & Q_SHADOW_L) {
                occ_l = trace_query(S, tp, R.ro, R.rd_l, R.t_l, true).hit != 0;
                ++n_queries;
            }
            if (pend & Q_SHADOW_E) {
                occ_e = trace_query(S, tp, R.ro, R.rd_e, inf, true).hit != 0;
                ++n_queries;
            }
            pend = 0;
        }
    }
}
int trace_block_size() { return TRACE_BLOCK; }

bool trace_needs_bvh(int n_tris, int n_spheres, int n_materials, int n_lights) {
    return n_tris > TRACE_LDS_TRIS || n_spheres > TRACE_FLAT_SPHERES || n_materials > TRACE_LDS_MATERIALS ||
           n_lights > TRACE_LDS_MATERIALS;
}

static size_t trace_lds_bytes(const DevScene& s, const DevBvh& b) {
    if (b.nodes) return size_t(BVH_STACK) * TRACE_BLOCK * sizeof(int);
    return std::max<size_t>(16, size_t(s.n_tris) * 10 * sizeof(float4) + size_t(s.n_materials) * sizeof(DevMaterial) +
                                    size_t(s.n_lights) * (sizeof(DevLight) + sizeof(float)));
}

// BVH scenes always take the depth-limited variant (max_depth 0 still means
// unbounded there); flat scenes take it only when a limit is set.
template <bool kLambert>
static const void* trace_kernel_m(const DevBvh& b, int max_depth) {
    if (b.nodes) return reinterpret_cast<const void*>(&k_trace<true, true, kLambert>);
    return max_depth > 0 ? reinterpret_cast<const void*>(&k_trace<false, true, kLambert>)
                         : reinterpret_cast<const void*>(&k_trace<false, false, kLambert>);
}
static const void* trace_kernel(const DevScene& s, const DevBvh& b, int max_depth) {
    return s.lambert_only ? trace_kernel_m<true>(b, max_depth) : trace_kernel_m<false>(b, max_depth);
}

int trace_blocks_per_cu(const DevScene& s, const DevBvh& b, int max_depth) {
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, trace_kernel(s, b, max_depth), TRACE_BLOCK,
                                                     trace_lds_bytes(s, b)) != hipSuccess || nb <= 0)
        nb = 1;
    return nb;
}

template <bool kLambert>
static void launch_trace_m(const TraceParams& p, int n_blocks, size_t lds, hipStream_t st) {
    if (p.bvh.nodes)
        hipLaunchKernelGGL((k_trace<true, true, kLambert>), dim3(n_blocks), dim3(TRACE_BLOCK), lds, st, p);
    else if (p.max_depth > 0)
        hipLaunchKernelGGL((k_trace<false, true, kLambert>), dim3(n_blocks), dim3(TRACE_BLOCK), lds, st, p);
    else
        hipLaunchKernelGGL((k_trace<false, false, kLambert>), dim3(n_blocks), dim3(TRACE_BLOCK), lds, st, p);
}

void launch_trace(const TraceParams& p, int n_blocks, hipStream_t st) {
    const size_t lds = trace_lds_bytes(p.scene, p.bvh);
    if (p.scene.lambert_only) launch_trace_m<true>(p, n_blocks, lds, st);
    else launch_trace_m<false>(p, n_blocks, lds, st);
}

// ---------------------------------------------------------------------------
// Wavefront trace of a flat scene. The path-regeneration kernel keeps each
// lane's whole path state in registers, so every loop iteration runs the
// shading code and the triangle loop at the occupancy the shading code
// allows, and each code block of an iteration with only the lanes that need
// it. Here the two halves are separate kernels over path slots held in HBM:
//
//  * k_wf_shade: one lane per slot calls path_step once (resolve the answered
//    queries, write a finished sample, hand out pixels, start a camera ray or
//    shade one bounce) and lists the slot's new queries;
//  * k_wf_isect: one lane per listed query runs the triangle loop, at the
//    register count of the loop alone.
//
// The same path_step, tapes and hand-out rules give the same records as the
// path-regeneration kernel. A fixed number of rounds is launched (the host
// cannot see when a pass ends); k_trace<..., kResume> then takes over the
// slots still running, so a pass ends however long its longest pixel is.
__global__ void k_wf_begin(WavefrontState w) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < w.slots) w.state[size_t(WF_ST) * w.slots + i] = uint32_t(ST_NEED);
    if (i < 2 * wf_groups(w.slots)) w.cursor[i] = 0;
    if (i < 2 * WF_MAX_ITERS * WF_LISTS) w.ctrl[i] = 0;
}

// True when some slot was still running after round `iter - 1` (or iter is
// the first round). Wave-uniform.
__device__ __forceinline__ bool wf_running(const WavefrontState& w, int iter, int lane) {
    if (iter == 0) return true;
    return ballot64(wf_live(w, iter - 1)[lane] != 0) != 0;
}

template <bool kDepth, bool kLambert>
__global__ void __launch_bounds__(TRACE_BLOCK) k_wf_shade(TraceParams p, int iter) {
    const int lane = threadIdx.x & 63;
    const int s = blockIdx.x * TRACE_BLOCK + threadIdx.x;
    const int g = s >> 6;
    if (!wf_running(p.wf, iter, lane)) return;   // the same for every wave of the launch
    DevScene S = p.scene;
    {
        extern __shared__ float4 s_dyn[];
        const DevScene& G = p.scene;
        const int n_mat4 = G.n_materials * int(sizeof(DevMaterial) / sizeof(float4));
        const int n_lig4 = G.n_lights * int(sizeof(DevLight) / sizeof(float4));
        float4* s_attr = s_dyn;
        float4* s_mat = s_attr + G.n_tris;
        float4* s_lig = s_mat + n_mat4;
        float* s_cdf = reinterpret_cast<float*>(s_lig + n_lig4);
        for (int k = threadIdx.x; k < G.n_tris; k += TRACE_BLOCK) s_attr[k] = G.tri_attr[k];
        for (int k = threadIdx.x; k < n_mat4; k += TRACE_BLOCK)
            s_mat[k] = reinterpret_cast<const float4*>(G.materials)[k];
        for (int k = threadIdx.x; k < n_lig4; k += TRACE_BLOCK) s_lig[k] = reinterpret_cast<const float4*>(G.lights)[k];
        for (int k = threadIdx.x; k < G.n_lights; k += TRACE_BLOCK) s_cdf[k] = G.light_cdf[k];
        __syncthreads();
        S.tri_attr = s_attr;
        S.materials = reinterpret_cast<const DevMaterial*>(s_mat);
        S.lights = reinterpret_cast<const DevLight*>(s_lig);
        S.light_cdf = s_cdf;
    }
    if ((g << 6) >= p.wf.slots) return;          // a wave past the last slot
    const v3 env = mk3(S.env[0], S.env[1], S.env[2]);
    const bool env_on = maxc3(env) > 0.0f;
    const int n_active = p.plan.totals[1];
    int c_next = p.wf.cursor[2 * g], c_end = p.wf.cursor[2 * g + 1];

    int st = ST_DEAD, q = 0, pend = 0;
    int pix = -1, left = 0;
    int64_t slot = 0;
    Draws dr;
    dr.tape = p.tape; dr.pos = 0; dr.limit = MT_N; dr.cache = make_uint4(0, 0, 0, 0);
    LaneRays R;
    R.ro = mk3(0, 0, 0); R.rd_ext = R.ro; R.rd_l = R.ro; R.rd_e = R.ro; R.t_l = 0.0f;
    v3 T = mk3(1, 1, 1), L = mk3(0, 0, 0), Ld = mk3(0, 0, 0);
    v3 sp = mk3(0, 0, 0), sn = mk3(0, 0, 0), swo = mk3(0, 0, 0);
    v3 cw_l = mk3(0, 0, 0), cw_e = mk3(0, 0, 0), cw_x = mk3(0, 0, 0);
    float bpdf = 0.0f;
    int smat = 0, depth = 0;
    uint32_t flags = 0, light_bits = 255u;
    HitRec h; h.hit = 0; h.sphere = 0; h.index = -1; h.t = 0.0f; h.sn = mk3(0, 0, 0);
    bool occ_l = false, occ_e = false;
    const bool valid = s < p.wf.slots;
    const WfRef W{p.wf, s};
    if (valid)
        wf_load<kDepth>(p, W, h, occ_l, occ_e, st, q, pend, pix, left, slot, dr, R, T, L, Ld, cw_l, cw_e, cw_x, bpdf,
                        smat, depth, flags, light_bits);
    path_step<kDepth, kLambert>(p, S, lane, env, env_on, c_next, c_end, n_active, h, occ_l, occ_e, st, q, pend, pix,
                                left, slot, dr, R, T, L, Ld, sp, sn, swo, cw_l, cw_e, cw_x, bpdf, smat, depth, flags,
                                light_bits);
    if (valid)
        wf_store<kDepth>(p, W, st, q, pix, left, slot, dr, R, T, L, Ld, cw_l, cw_e, cw_x, bpdf, smat, depth, flags,
                         light_bits);
    if (lane == 0) { p.wf.cursor[2 * g] = c_next; p.wf.cursor[2 * g + 1] = c_end; }

    // list the queries: this group's sub-list, one counter update per wave
    const int list = g % WF_LISTS;
    const uint64_t below = (1ull << lane) - 1ull;
    const uint64_t mx = ballot64(pend & Q_EXT), ml = ballot64(pend & Q_SHADOW_L), me = ballot64(pend & Q_SHADOW_E);
    const int nx = __popcll(mx), nl = __popcll(ml), n = nx + nl + __popcll(me);
    if (n) {
        int base = 0;
        if (lane == 0) base = atomicAdd(wf_count(p.wf, iter) + list, n);
        base = __shfl(base, 0);
        int* out = p.wf.queue + size_t(list) * wf_list_cap(p.wf.slots) + base;
        if (pend & Q_EXT) out[__popcll(mx & below)] = s * 4 + 1;
        if (pend & Q_SHADOW_L) out[nx + __popcll(ml & below)] = s * 4 + 2;
        if (pend & Q_SHADOW_E) out[nx + nl + __popcll(me & below)] = s * 4 + 3;
    }
    if (lane == 0 && ballot64(st != ST_DEAD) != 0) wf_live(p.wf, iter)[list] = 1;
}

constexpr int WF_ISECT_BLOCK = 256;
__global__ void __launch_bounds__(WF_ISECT_BLOCK) k_wf_isect(TraceParams p, int iter) {
    __shared__ int s_count[WF_LISTS];
    extern __shared__ float4 s_dyn[];
    const int* cnt = wf_count(p.wf, iter);
    if (threadIdx.x < WF_LISTS) s_count[threadIdx.x] = cnt[threadIdx.x];
    const DevScene& S = p.scene;
    const int n_perm = S.n_tris * 9;
    for (int k = threadIdx.x; k < n_perm; k += WF_ISECT_BLOCK) s_dyn[k] = S.tri_perm[k];
    __syncthreads();
    const int cap = wf_list_cap(p.wf.slots);
    const float inf = __builtin_inff();
    uint32_t n_queries = 0;
    const int64_t total = int64_t(WF_LISTS) * cap;
    for (int64_t w = int64_t(blockIdx.x) * WF_ISECT_BLOCK + threadIdx.x; w < total;
         w += int64_t(gridDim.x) * WF_ISECT_BLOCK) {
        const int list = int(w / cap), e = int(w - int64_t(list) * cap);
        if (e >= s_count[list]) continue;
        const int code = p.wf.queue[w];
        const int kind = code & 3;
        const WfRef W{p.wf, code >> 2};
        const v3 ro = W.f3(WF_RO);
        if (kind == 1) {
            const HitRec h = trace_query(S, s_dyn, ro, W.f3(WF_RDX), inf, false);
            W.put(WF_HCODE, h.hit ? uint32_t(h.index * 2 + h.sphere) : 0xffffffffu);
            W.putf(WF_HT, h.t);
            if (h.sphere) W.put3(WF_HSN, h.sn);
        } else if (kind == 2) {
            W.put(WF_OCCL, trace_query(S, s_dyn, ro, W.f3(WF_RDL), W.f(WF_TL), true).hit ? 1u : 0u);
        } else {
            W.put(WF_OCCE, trace_query(S, s_dyn, ro, W.f3(WF_RDE), inf, true).hit ? 1u : 0u);
        }
        ++n_queries;
    }
    if (p.query_count) {
        unsigned long long tot = n_queries;
        for (int o = 32; o > 0; o >>= 1) tot += __shfl_xor(tot, o);
        if ((threadIdx.x & 63) == 0 && tot) atomicAdd(p.query_count, tot);
    }
}

static size_t wf_shade_lds_bytes(const DevScene& s) {
    return std::max<size_t>(16, size_t(s.n_tris) * sizeof(float4) + size_t(s.n_materials) * sizeof(DevMaterial) +
                                    size_t(s.n_lights) * (sizeof(DevLight) + sizeof(float)));
}

template <bool kLambert>
static void launch_wavefront_m(const TraceParams& p, int iterations, int finish_blocks, hipStream_t st) {
    const WavefrontState& w = p.wf;
    const int begin_n = std::max(w.slots, 2 * WF_MAX_ITERS * WF_LISTS);
    hipLaunchKernelGGL(k_wf_begin, dim3((begin_n + 255) / 256), dim3(256), 0, st, w);
    const int shade_blocks = (w.slots + TRACE_BLOCK - 1) / TRACE_BLOCK;
    const size_t shade_lds = wf_shade_lds_bytes(p.scene);
    const size_t isect_lds = std::max<size_t>(16, size_t(p.scene.n_tris) * 9 * sizeof(float4));
    static int isect_blocks = [] {
        int nb = 0, dev = 0, cus = 0;
        hipGetDevice(&dev);
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, reinterpret_cast<const void*>(&k_wf_isect),
                                                         WF_ISECT_BLOCK, 9 * TRACE_LDS_TRIS * sizeof(float4)) !=
                hipSuccess || nb <= 0)
            nb = 1;
        return nb * std::max(1, cus);
    }();
    const int n_iter = std::min(iterations, WF_MAX_ITERS);
    for (int i = 0; i < n_iter; ++i) {
        if (p.max_depth > 0)
            hipLaunchKernelGGL((k_wf_shade<true, kLambert>), dim3(shade_blocks), dim3(TRACE_BLOCK), shade_lds, st, p, i);
        else
            hipLaunchKernelGGL((k_wf_shade<false, kLambert>), dim3(shade_blocks), dim3(TRACE_BLOCK), shade_lds, st, p,
                               i);
        hipLaunchKernelGGL(k_wf_isect, dim3(isect_blocks), dim3(WF_ISECT_BLOCK), isect_lds, st, p, i);
    }
    const int fb = std::max(finish_blocks, shade_blocks);
    const size_t lds = trace_lds_bytes(p.scene, p.bvh);
    if (p.max_depth > 0)
        hipLaunchKernelGGL((k_trace<false, true, kLambert, true>), dim3(fb), dim3(TRACE_BLOCK), lds, st, p);
    else
        hipLaunchKernelGGL((k_trace<false, false, kLambert, true>), dim3(fb), dim3(TRACE_BLOCK), lds, st, p);
}

void launch_trace_wavefront(const TraceParams& p, int iterations, int finish_blocks, hipStream_t st) {
    if (p.scene.lambert_only) launch_wavefront_m<true>(p, iterations, finish_blocks, st);
    else launch_wavefront_m<false>(p, iterations, finish_blocks, st);
}

// ---------------------------------------------------------------------------
// Pass plans. A uniform pass samples every pixel `count` times; an adaptive
// pass samples counts[p] times. offsets[p] is the pixel's first sample slot.
__global__ void k_plan_uniform(int n, int count, int* counts, int* offsets, int* active, int* totals) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) { totals[0] = n * count; totals[1] = count > 0 ? n : 0; totals[2] = 0; }
    if (i >= n) return;
    counts[i] = count;
    offsets[i] = i * count;
    active[i] = i;
}
void launch_plan_uniform(int n, int count, int* counts, int* offsets, int* active, int* totals, hipStream_t st) {
    hipLaunchKernelGGL(k_plan_uniform, dim3((n + 255) / 256), dim3(256), 0, st, n, count, counts, offsets, active,
                       totals);
}
__global__ void k_fill_i32(int* p, int v, int64_t n) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}
void launch_fill_i32(int* p, int v, int64_t n, hipStream_t st) {
    hipLaunchKernelGGL(k_fill_i32, dim3((n + 255) / 256), dim3(256), 0, st, p, v, n);
}

// Exclusive scan of (active flag << 32 | count) in three steps: per-block
// inclusive sums, one block scanning the block totals, then the fix-up that
// writes offsets and the compacted active list.
constexpr int SCAN_BLOCK = 1024;

__device__ __forceinline__ int64_t block_incl_scan(int64_t v, int64_t* sh) {
    const int t = threadIdx.x;
    sh[t] = v;
    __syncthreads();
    for (int o = 1; o < SCAN_BLOCK; o <<= 1) {
        const int64_t add = t >= o ? sh[t - o] : 0;
        __syncthreads();
        sh[t] += add;
        __syncthreads();
    }
    return sh[t];
}

__global__ void __launch_bounds__(SCAN_BLOCK) k_scan_blocks(int n, const int* counts, int64_t* block_sums) {
    __shared__ int64_t sh[SCAN_BLOCK];
    const int i = blockIdx.x * SCAN_BLOCK + threadIdx.x;
    int64_t v = 0;
    if (i < n) { const int c = counts[i]; v = (int64_t(c > 0 ? 1 : 0) << 32) | int64_t(c); }
    const int64_t s = block_incl_scan(v, sh);
    if (threadIdx.x == SCAN_BLOCK - 1) block_sums[blockIdx.x] = s;
}
__global__ void __launch_bounds__(SCAN_BLOCK) k_scan_totals(int nb, int64_t* block_sums, int* totals,
                                                            int64_t capacity) {
    __shared__ int64_t sh[SCAN_BLOCK];
    int64_t carry = 0;
    for (int base = 0; base < nb; base += SCAN_BLOCK) {
        const int i = base + threadIdx.x;
        const int64_t v = i < nb ? block_sums[i] : 0;
        const int64_t s = block_incl_scan(v, sh);
        if (i < nb) block_sums[i] = carry + s - v;  // exclusive
        __syncthreads();
        carry += sh[SCAN_BLOCK - 1];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const int64_t samples = carry & 0xffffffffll;
        const bool over = samples > capacity;
        totals[0] = int(samples);
        totals[1] = over ? 0 : int(carry >> 32);
        totals[2] = over ? 1 : 0;
    }
}
__global__ void __launch_bounds__(SCAN_BLOCK) k_scan_apply(int n, const int* counts, const int64_t* block_sums,
                                                           int* offsets, int* active) {
    __shared__ int64_t sh[SCAN_BLOCK];
    const int i = blockIdx.x * SCAN_BLOCK + threadIdx.x;
    int64_t v = 0;
    int c = 0;
    if (i < n) { c = counts[i]; v = (int64_t(c > 0 ? 1 : 0) << 32) | int64_t(c); }
    const int64_t s = block_incl_scan(v, sh) - v + block_sums[blockIdx.x];
    if (i < n) {
        offsets[i] = int(s & 0xffffffffll);
        if (c > 0) active[int(s >> 32)] = i;
    }
}
int64_t plan_scan_scratch_elems(int n) { return (n + SCAN_BLOCK - 1) / SCAN_BLOCK + 1; }

// Tile order. The trace hands out the active list in order, 64 pixels to a
// wave at a time; in row order those are a 64x1 strip, in tile order an 8x8
// tile, whose camera rays, first hits and light directions lie closer
// together (the BVH walk diverges less across the lanes, the shading less
// across materials). This rewrites the active list of a plan in tile order
// (8x8 tiles row by row, pixels row by row within a tile): the same pixels,
// counts, offsets and records, handed out in another order.
//
// With long_from > 0 the pixels of at least long_from samples go first (in tile
// order), then the others. A pixel's samples are one chain on one lane and an
// adaptive pass gives some pixels up to eight times the mean, so those chains
// start at once; the trace's wave priority then runs them ahead of the rest.
// Both classes are counted in one scan: class 0 in the low 32 bits of the
// int64 prefix, class 1 in the high 32 bits (each count stays below 2^26).
constexpr int PLAN_TILE = 8;
__host__ __device__ inline int64_t plan_tile_positions(int W, int H) {
    return int64_t((W + PLAN_TILE - 1) / PLAN_TILE) * int64_t((H + PLAN_TILE - 1) / PLAN_TILE) * PLAN_TILE *
           PLAN_TILE;
}
__device__ __forceinline__ int tile_pixel(int64_t k, int W, int H) {
    const int tiles_x = (W + PLAN_TILE - 1) / PLAN_TILE;
    const int64_t t = k / (PLAN_TILE * PLAN_TILE);
    const int w = int(k % (PLAN_TILE * PLAN_TILE));
    const int x = int(t % tiles_x) * PLAN_TILE + w % PLAN_TILE, y = int(t / tiles_x) * PLAN_TILE + w / PLAN_TILE;
    return (x < W && y < H) ? y * W + x : -1;
}
__device__ __forceinline__ int64_t tord_class_bit(int c, int long_from) {
    if (c <= 0) return 0;
    return (long_from > 0 && c >= long_from) ? int64_t(1) : (int64_t(1) << 32);
}
__global__ void __launch_bounds__(SCAN_BLOCK) k_tord_blocks(int W, int H, int64_t K, const int* counts, int long_from,
                                                            int64_t* block_sums) {
    __shared__ int64_t sh[SCAN_BLOCK];
    const int64_t k = int64_t(blockIdx.x) * SCAN_BLOCK + threadIdx.x;
    const int p = k < K ? tile_pixel(k, W, H) : -1;
    const int64_t v = p >= 0 ? tord_class_bit(counts[p], long_from) : 0;
    const int64_t s = block_incl_scan(v, sh);
    if (threadIdx.x == SCAN_BLOCK - 1) block_sums[blockIdx.x] = s;
}
// Exclusive prefix of the block sums in place; block_sums[nb] gets the total.
__global__ void __launch_bounds__(SCAN_BLOCK) k_tord_totals(int nb, int64_t* block_sums) {
    __shared__ int64_t sh[SCAN_BLOCK];
    int64_t carry = 0;
    for (int base = 0; base < nb; base += SCAN_BLOCK) {
        const int i = base + threadIdx.x;
        const int64_t v = i < nb ? block_sums[i] : 0;
        const int64_t s = block_incl_scan(v, sh);
        if (i < nb) block_sums[i] = carry + s - v;  // exclusive
        __syncthreads();
        carry += sh[SCAN_BLOCK - 1];
        __syncthreads();
    }
    if (threadIdx.x == 0) block_sums[nb] = carry;
}
__global__ void __launch_bounds__(SCAN_BLOCK) k_tord_apply(int W, int H, int64_t K, const int* counts, int long_from,
                                                           int nb, const int64_t* block_sums, int* active) {
    __shared__ int64_t sh[SCAN_BLOCK];
    const int64_t k = int64_t(blockIdx.x) * SCAN_BLOCK + threadIdx.x;
    const int p = k < K ? tile_pixel(k, W, H) : -1;
    const int64_t v = p >= 0 ? tord_class_bit(counts[p], long_from) : 0;
    const int64_t s = block_incl_scan(v, sh) - v + block_sums[blockIdx.x];
    if (v == 1) active[s & 0xffffffffll] = p;
    else if (v != 0) active[(block_sums[nb] & 0xffffffffll) + (s >> 32)] = p;
}
int64_t plan_tile_scratch_elems(int W, int H) { return (plan_tile_positions(W, H) + SCAN_BLOCK - 1) / SCAN_BLOCK + 1; }
void launch_plan_tile_order(int W, int H, const int* counts, int* active, int64_t* scratch, int long_from,
                            hipStream_t st) {
    const int64_t K = plan_tile_positions(W, H);
    const int nb = int((K + SCAN_BLOCK - 1) / SCAN_BLOCK);
    hipLaunchKernelGGL(k_tord_blocks, dim3(nb), dim3(SCAN_BLOCK), 0, st, W, H, K, counts, long_from, scratch);
    hipLaunchKernelGGL(k_tord_totals, dim3(1), dim3(SCAN_BLOCK), 0, st, nb, scratch);
    hipLaunchKernelGGL(k_tord_apply, dim3(nb), dim3(SCAN_BLOCK), 0, st, W, H, K, counts, long_from, nb, scratch,
                       active);
}
void launch_plan_scan(int n, const int* counts, int* offsets, int* active, int* totals, int64_t* scratch,
                      int64_t capacity, hipStream_t st) {
    const int nb = (n + SCAN_BLOCK - 1) / SCAN_BLOCK;
    hipLaunchKernelGGL(k_scan_blocks, dim3(nb), dim3(SCAN_BLOCK), 0, st, n, counts, scratch);
    hipLaunchKernelGGL(k_scan_totals, dim3(1), dim3(SCAN_BLOCK), 0, st, nb, scratch, totals, capacity);
    hipLaunchKernelGGL(k_scan_apply, dim3(nb), dim3(SCAN_BLOCK), 0, st, n, counts, scratch, offsets, active);
}

// ---------------------------------------------------------------------------
// Fold a pass's samples into the per-pixel accumulators in sample order
// (main.cpp:572-576, 707-717; core/stats.h:25-41). Features are summed for every
// sample with a primary hit, statistics and radiance only for finite samples.
// The records of FOLD_AHEAD samples are loaded before the first of them is
// folded (all four planes, hit or not), so a lane waits on memory once per
// group rather than twice per sample; the folding order is unchanged.
constexpr int FOLD_AHEAD = 2;
__global__ void __launch_bounds__(256) k_fold(Batch b, int statmc) {
    const int64_t item = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (item >= b.total[BK_PIXEL]) return;
    int64_t la;
    const ImageDesc& d = b.img[batch_find(b, BK_PIXEL, item, la)];
    const PassPlan& plan = d.plan;
    const Records& rec = d.rec;
    const PixelBuffers& pb = d.pb;
    // one lane per active pixel, in the active list's order (walking all
    // pixels and skipping the inactive ones measured slower: the C3 batch
    // 447 -> 476 ms)
    const int a = int(la);
    if (a >= plan.totals[1]) return;
    const int i = plan.active[a];
    const int N = plan.n_pixels;
    const int cnt = plan.counts[i];
    const int64_t off = plan.offsets[i];
    float rr = pb.radiance[i], rg = pb.radiance[N + i], rb = pb.radiance[2 * N + i];
    float ar = pb.albedo[i], ag = pb.albedo[N + i], ab = pb.albedo[2 * N + i];
    float nx = pb.normal[i], ny = pb.normal[N + i], nz = pb.normal[2 * N + i];
    float px = pb.world_pos[i], py = pb.world_pos[N + i], pz = pb.world_pos[2 * N + i];
    float dep = pb.depth[i];
    int hits = pb.hit_count[i], scount = pb.sample_count[i];
    float mean = 0, m2 = 0;
    int n = 0;
    v3 cm = mk3(0, 0, 0), sm = mk3(0, 0, 0), s2 = mk3(0, 0, 0);
    if (statmc) {
        mean = pb.lum_mean[i]; m2 = pb.lum_m2[i]; n = pb.n[i];
        cm = mk3(pb.color_mean[i], pb.color_mean[N + i], pb.color_mean[2 * N + i]);
        sm = mk3(pb.sqrt_mean[i], pb.sqrt_mean[N + i], pb.sqrt_mean[2 * N + i]);
        s2 = mk3(pb.sqrt_m2[i], pb.sqrt_m2[N + i], pb.sqrt_m2[2 * N + i]);
    }
    for (int s0 = 0; s0 < cnt; s0 += FOLD_AHEAD) {
        float4 q0[FOLD_AHEAD], q1[FOLD_AHEAD], q2[FOLD_AHEAD], q3[FOLD_AHEAD];
#pragma unroll
        for (int k = 0; k < FOLD_AHEAD; ++k) {
            if (s0 + k < cnt) {
                const int64_t o = off + s0 + k;
                q0[k] = rec.r0[o]; q1[k] = rec.r1[o]; q2[k] = rec.r2[o]; q3[k] = rec.r3[o];
            }
        }
#pragma unroll
        for (int k = 0; k < FOLD_AHEAD; ++k) {
            if (s0 + k >= cnt) break;
            const float4 r0 = q0[k];
            const uint32_t fl = __float_as_uint(r0.w);
            if (fl & REC_HIT) {
                const float4 r1 = q1[k], r2 = q2[k], r3 = q3[k];
                ++hits;
                nx += r1.x; ny += r1.y; nz += r1.z;
                ar += r2.x; ag += r2.y; ab += r2.z;
                px += r3.x; py += r3.y; pz += r3.z;
                dep += r1.w;
            }
            if (!(fl & REC_FINITE)) continue;
            const v3 c = mk3(r0.x, r0.y, r0.z);
            if (statmc) {
                const float lum = luminance3(c);
                const float d = lum - mean;
                ++n;
                mean += d / float(n);
                const float d2 = lum - mean;
                m2 += d * d2;
                cm = cm + (c - cm) / float(n);
                const v3 sc = mk3(sqrtf(fmaxf(0.0f, c.x)), sqrtf(fmaxf(0.0f, c.y)), sqrtf(fmaxf(0.0f, c.z)));
                const v3 sd = sc - sm;
                sm = sm + sd / float(n);
                s2 = s2 + vmul(sd, sc - sm);
            }
            rr += c.x; rg += c.y; rb += c.z;
            ++scount;
        }
    }
    pb.radiance[i] = rr; pb.radiance[N + i] = rg; pb.radiance[2 * N + i] = rb;
    pb.albedo[i] = ar; pb.albedo[N + i] = ag; pb.albedo[2 * N + i] = ab;
    pb.normal[i] = nx; pb.normal[N + i] = ny; pb.normal[2 * N + i] = nz;
    pb.world_pos[i] = px; pb.world_pos[N + i] = py; pb.world_pos[2 * N + i] = pz;
    pb.depth[i] = dep;
    pb.hit_count[i] = hits;
    pb.sample_count[i] = scount;
    if (statmc) {
        pb.lum_mean[i] = mean; pb.lum_m2[i] = m2; pb.n[i] = n;
        pb.color_mean[i] = cm.x; pb.color_mean[N + i] = cm.y; pb.color_mean[2 * N + i] = cm.z;
        pb.sqrt_mean[i] = sm.x; pb.sqrt_mean[N + i] = sm.y; pb.sqrt_mean[2 * N + i] = sm.z;
        pb.sqrt_m2[i] = s2.x; pb.sqrt_m2[N + i] = s2.y; pb.sqrt_m2[2 * N + i] = s2.z;
    }
}

// Each lane reads its pixel's samples one after another, 16 bytes per plane
// per sample, so a lane comes back to each cache line over several samples.
// With every wave of a full CU in flight those lines are evicted from L2
// between the visits; the launch asks for LDS it does not use so that at most
// two blocks (8 waves) are resident per CU, which keeps them (the C3 batch of
// 32 images: 470 -> 444 ms; 16 waves per CU measured no better than 32).
constexpr size_t FOLD_RESIDENCY_LDS = 80 * 1024;
void launch_fold(const Batch& b, int statmc, hipStream_t st) {
    hipLaunchKernelGGL(k_fold, dim3((b.total[BK_PIXEL] + 255) / 256), dim3(256), FOLD_RESIDENCY_LDS, st, b, statmc);
}

// ---------------------------------------------------------------------------
// core/buffers.h:152-167: averages over samples and primary hits.
__global__ void k_finalize(int n, PixelBuffers pb, float* orad, float* oalb, float* onrm, float* opos, float* odep) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float s = float(max(1, pb.sample_count[i])), h = float(max(1, pb.hit_count[i]));
    for (int c = 0; c < 3; ++c) {
        orad[c * n + i] = pb.radiance[c * n + i] / s;
        oalb[c * n + i] = pb.albedo[c * n + i] / h;
        opos[c * n + i] = pb.world_pos[c * n + i] / h;
    }
    v3 nn = mk3(pb.normal[i] / h, pb.normal[n + i] / h, pb.normal[2 * n + i] / h);
    if (sqnorm3(nn) > RPD_EPS_SMALL) nn = normalized3(nn);
    onrm[i] = nn.x; onrm[n + i] = nn.y; onrm[2 * n + i] = nn.z;
    odep[i] = pb.depth[i] / h;
}
void launch_finalize(int n, const PixelBuffers& pb, float* orad, float* oalb, float* onrm, float* opos, float* odep,
                     hipStream_t st) {
    hipLaunchKernelGGL(k_finalize, dim3((n + 255) / 256), dim3(256), 0, st, n, pb, orad, oalb, onrm, opos, odep);
}

}  // namespace rpd
