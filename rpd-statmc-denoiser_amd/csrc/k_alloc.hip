// Adaptive sample allocation (path_tracer/main.cpp:1273-1446).
//
// The reference allocates with sequential FP64 arithmetic: std::accumulate of
// the importance field, an in-order clip-and-redistribute sweep, floors, and a
// systematic residual draw against a running FP64 cumulative sum. Every count
// is reproduced bit for bit here without running those loops serially:
//
//  * Sequential FP64 sums of non-negative floats. S_k = fl(S_{k-1} + x_k) only
//    depends on S_{k-1} through its binade and, on an exact tie, its parity.
//    Inside one binade [2^E, 2^(E+1)) every double is a multiple of
//    u = 2^(E-52), so a step that stays in the binade and is not a tie adds
//    exactly round(x_k / u) units of u. Steps are classified in parallel from an
//    approximate prefix (with a margin far above its rounding error); the
//    integer unit counts of each run of such steps are summed exactly, and one
//    lane replays only the remaining steps (binade changes, ties, the first
//    step), in order, with the reference's own double additions.
//  * Clipping: the first pixel that exceeds the cap under the initial state is
//    found by a grid-wide search; sweeps continue serially only if one exists.
//  * Residual draw: sample k lands on the first pixel whose cumulative residual
//    reaches target_k = ((k + offset) / left) * rsum, so pixel c receives
//    #{k : target_k <= P_c} - #{k : target_k <= P_{c-1}}, counted per pixel
//    from the exact prefix sums P.
#include "kernels.h"

namespace rpd {

// main.cpp:1305-1328: variance source and sqrt importance
__global__ void k_importance(int n, PixelBuffers pb, const float* vden, float smax, float* imp, float* dbg_src) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int cnt = pb.n[i];
    float raw = cnt > 1 ? pb.lum_m2[i] / float(cnt - 1) : 0.0f;
    if (!isfinite(raw)) raw = RPD_EPS_SMALL;
    raw = fmaxf(RPD_EPS_SMALL, raw);
    float d = vden[i];
    if (!isfinite(d) || d <= 0.0f) d = raw;
    const float src = clampf(fmaxf(raw, d), RPD_EPS_SMALL, smax * smax);
    if (dbg_src) dbg_src[i] = src;
    imp[i] = sqrtf(src);
}

// main.cpp:1331-1368: edge-aware smoothing
__global__ void __launch_bounds__(SMOOTH_TW * SMOOTH_TH) k_importance_smooth(int W, int H, int R, PixelBuffers pb,
                                                                             Features f, const float* src,
                                                                             float* dst, int staged) {
    extern __shared__ float4 s_tile[];
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
    double acc = 0.0, ws = 0.0;
    for (int dy = -R; dy <= R; ++dy) {
        const int yy = y + dy;
        if (yy < 0 || yy >= H) continue;
        for (int dx = -R; dx <= R; ++dx) {
            const int xx = x + dx;
            if (xx < 0 || xx >= W) continue;
            float4 gq;
            float vq;
            int hq;
            S.get(xx, yy, gq, vq, hq);
            if (hq == 0) continue;
            const float sw = 1.0f / (1.0f + float(dx * dx + dy * dy));
            const float nw = clampf((dot3(ni, mk3(gq.x, gq.y, gq.z)) + 1.0f) * 0.5f, 0.0f, 1.0f);
            const float dw = expf_cr(-4.0f * relative_depth_delta(di, gq.w));
            const double w = double(sw) * double(nw) * double(dw);
            acc += w * double(vq);
            ws += w;
        }
    }
    dst[i] = (ws > double(RPD_EPS_SMALL)) ? float(acc / ws) : vsrc;
}

// ---------------------------------------------------------------------------
// Generic inclusive scans (three launches: block scans, block totals, fix-up).
constexpr int SB = 1024;

template <typename T>
__device__ __forceinline__ T block_scan_incl(T v, T* sh) {
    const int t = threadIdx.x;
    sh[t] = v;
    __syncthreads();
    for (int o = 1; o < SB; o <<= 1) {
        const T add = t >= o ? sh[t - o] : T(0);
        __syncthreads();
        sh[t] = sh[t] + add;
        __syncthreads();
    }
    return sh[t];
}
template <typename T>
__global__ void __launch_bounds__(SB) k_scan_a(int n, T* data, T* bsum) {
    __shared__ T sh[SB];
    const int i = blockIdx.x * SB + threadIdx.x;
    const T v = i < n ? data[i] : T(0);
    const T s = block_scan_incl(v, sh);
    if (i < n) data[i] = s;
    if (threadIdx.x == SB - 1) bsum[blockIdx.x] = s;
}
template <typename T>
__global__ void __launch_bounds__(SB) k_scan_b(int nb, T* bsum) {  // exclusive scan of block totals
    __shared__ T sh[SB];
    T carry = T(0);
    for (int base = 0; base < nb; base += SB) {
        const int i = base + threadIdx.x;
        const T v = i < nb ? bsum[i] : T(0);
        const T s = block_scan_incl(v, sh);
        if (i < nb) bsum[i] = carry + (s - v);
        __syncthreads();
        carry = carry + sh[SB - 1];
        __syncthreads();
    }
}
template <typename T>
__global__ void k_scan_c(int n, T* data, const T* bsum) {
    const int i = blockIdx.x * SB + threadIdx.x;
    if (i < n && blockIdx.x > 0) data[i] = data[i] + bsum[blockIdx.x];
}
template <typename T>
void scan_inclusive(int n, T* data, T* bsum, hipStream_t st) {
    const int nb = (n + SB - 1) / SB;
    hipLaunchKernelGGL(k_scan_a<T>, dim3(nb), dim3(SB), 0, st, n, data, bsum);
    hipLaunchKernelGGL(k_scan_b<T>, dim3(1), dim3(SB), 0, st, nb, bsum);
    hipLaunchKernelGGL(k_scan_c<T>, dim3(nb), dim3(SB), 0, st, n, data, bsum);
}

// ---------------------------------------------------------------------------
// Exact sequential FP64 sum of non-negative floats.
struct SeqSum {
    int n;
    const float* x;
    double* approx;        // n: approximate inclusive prefix, then (optionally) the exact prefix
    uint64_t* units;       // n: unit counts of simple steps, then their inclusive scan
    int* serial_pos;       // n: serial flag, then its inclusive scan
    int* serial_list;      // positions of serial steps, in order
    int* binade;           // n: binade exponent of simple steps
    double* seg_add;       // per serial step: exact sum of the simple run before it (+1 for the tail)
    double* s_at;          // per serial step: S right after it
    double* total;         // 1: the final sum
    int* n_serial;         // 1
    double* bsum_d;
    uint64_t* bsum_u;
    int* bsum_i;
};

__device__ __forceinline__ int binade_of(double v) {  // floor(log2 v) for v > 0
    int e;
    frexp(v, &e);
    return e - 1;
}

__global__ void k_seq_init(SeqSum s) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < s.n) s.approx[i] = double(s.x[i]);
}

// The approximate prefix (any summation order) is within (n + 64) * 2^-52 of
// the sequential one in relative terms, below 2^-24 for up to 2^26 terms.
#define SEQ_MARGIN 1e-7

__global__ void k_seq_classify(SeqSum s) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= s.n) return;
    const float xf = s.x[k];
    bool serial = (k == 0);
    uint64_t q = 0;
    int E = 0;
    if (!serial) {
        const double lo_prev = s.approx[k - 1] * (1.0 - SEQ_MARGIN);
        const double hi_cur = s.approx[k] * (1.0 + SEQ_MARGIN);
        if (s.approx[k - 1] == 0.0) {
            // the terms are non-negative, so a zero prefix means every term so far is zero and
            // the sum is exactly +0; adding another zero keeps it (a simple step of no units)
            serial = xf != 0.0f;
        } else if (!(lo_prev > 0.0)) {
            serial = true;                   // running sum may still be zero
        } else {
            E = binade_of(lo_prev);
            const double hi_prev = s.approx[k - 1] * (1.0 + SEQ_MARGIN);
            // both S_{k-1} (whole interval) and S_k certainly in binade E; this also makes
            // consecutive simple steps share E even where the approximate prefix is not monotone
            if (binade_of(hi_cur) != E || binade_of(hi_prev) != E) {
                serial = true;               // may cross a binade boundary
            } else if (xf != 0.0f) {
                // x = M * 2^(fe-24) with integer M < 2^24; in units u = 2^(E-52): x/u = M * 2^(fe-24-E+52)
                int fe;
                const float fm = frexpf(xf, &fe);
                const uint32_t M = uint32_t(ldexpf(fm, 24));
                const int sh = fe - 24 - E + 52;
                if (sh >= 0) {
                    q = uint64_t(M) << sh;                            // < 2^53 since x < 2^(E+1)
                } else if (sh > -40) {
                    const int r = -sh;
                    const uint64_t half = 1ull << (r - 1);
                    const uint64_t rem = uint64_t(M) & ((1ull << r) - 1ull);
                    if (rem == half) serial = true;                   // exact tie: parity decides
                    else q = (uint64_t(M) + half) >> r;
                }                                                     // else far below half a unit: q = 0
            }
        }
    }
    s.units[k] = serial ? 0ull : q;
    s.serial_pos[k] = serial ? 1 : 0;
    s.binade[k] = E;
}

__global__ void k_seq_compact(SeqSum s) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= s.n) return;
    const int c = s.serial_pos[k];
    const bool is_serial = (k == 0) ? c == 1 : c != s.serial_pos[k - 1];
    if (is_serial) s.serial_list[c - 1] = k;
    if (k == s.n - 1) *s.n_serial = c;
}

// Run j (0 <= j <= m): the simple steps just before serial step j; run m is the tail.
__global__ void k_seq_runs(SeqSum s) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    const int m = *s.n_serial;
    if (j > m) return;
    const int start = (j == 0) ? 0 : s.serial_list[j - 1] + 1;
    const int end = (j == m) ? s.n : s.serial_list[j];      // exclusive
    double add = 0.0;
    if (end > start) {
        const uint64_t Q = s.units[end - 1] - s.units[start - 1];   // units are zero at serial steps
        add = ldexp(double(Q), s.binade[start] - 52);              // exact: Q < 2^53
    }
    s.seg_add[j] = add;
}

// The serial chain: one lane adds each run (exactly), then the serial step
// itself with the reference's rounded addition. Other lanes stage inputs.
constexpr int SEQ_STAGE = 2048;
__global__ void __launch_bounds__(SB) k_seq_chain(SeqSum s) {
    __shared__ double st_add[2][SEQ_STAGE];
    __shared__ float st_x[2][SEQ_STAGE];
    __shared__ double st_out[SEQ_STAGE];
    const int t = threadIdx.x;
    const int m = *s.n_serial;
    double S = 0.0;
    const int nch = (m + SEQ_STAGE - 1) / SEQ_STAGE;
    for (int k = t; k < SEQ_STAGE; k += SB)
        if (k < m) { st_add[0][k] = s.seg_add[k]; st_x[0][k] = s.x[s.serial_list[k]]; }
    __syncthreads();
    for (int c = 0; c < nch; ++c) {
        const int base = c * SEQ_STAGE, cnt = min(SEQ_STAGE, m - base), b = c & 1;
        if (t == 0) {
            for (int k = 0; k < cnt; ++k) {
                S = S + st_add[b][k];
                S = S + double(st_x[b][k]);
                st_out[k] = S;
            }
        } else if (c + 1 < nch) {
            const int nb = base + SEQ_STAGE;
            for (int k = t - 1; k < SEQ_STAGE; k += SB - 1)
                if (nb + k < m) { st_add[b ^ 1][k] = s.seg_add[nb + k]; st_x[b ^ 1][k] = s.x[s.serial_list[nb + k]]; }
        }
        __syncthreads();
        for (int k = t; k < cnt; k += SB) s.s_at[base + k] = st_out[k];
        __syncthreads();
    }
    if (t == 0) *s.total = S + s.seg_add[m];
}

// exact prefix S_k for every k (into `approx`)
__global__ void k_seq_prefix(SeqSum s) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= s.n) return;
    const int c = s.serial_pos[k];                 // serial steps at or before k
    const int last = s.serial_list[c - 1];
    if (last == k) { s.approx[k] = s.s_at[c - 1]; return; }
    const uint64_t Q = s.units[k] - s.units[last];
    s.approx[k] = s.s_at[c - 1] + ldexp(double(Q), s.binade[k] - 52);
}

void seq_sum_launch(const SeqSum& s, bool want_prefix, hipStream_t st) {
    const int n = s.n;
    const dim3 g((n + 255) / 256), b(256);
    hipLaunchKernelGGL(k_seq_init, g, b, 0, st, s);
    scan_inclusive<double>(n, s.approx, s.bsum_d, st);
    hipLaunchKernelGGL(k_seq_classify, g, b, 0, st, s);
    scan_inclusive<uint64_t>(n, s.units, s.bsum_u, st);
    scan_inclusive<int>(n, s.serial_pos, s.bsum_i, st);
    hipLaunchKernelGGL(k_seq_compact, g, b, 0, st, s);
    hipLaunchKernelGGL(k_seq_runs, dim3((n + 1 + 255) / 256), b, 0, st, s);
    hipLaunchKernelGGL(k_seq_chain, dim3(1), dim3(SB), 0, st, s);
    if (want_prefix) hipLaunchKernelGGL(k_seq_prefix, g, b, 0, st, s);
}

// ---------------------------------------------------------------------------
// Allocation state shared by the kernels below (device memory).
struct AllocState {
    double rem_imp;    // remaining importance after clipping
    int remaining;     // remaining samples after clipping
    int first_clip;    // first clip of the first sweep (INT_MAX: none)
    int final_assign;  // 1 if the sweep loop ended by assigning shares
    int assigned;      // sum of floors
    int active;        // 0: allocation ended early (counts = floors)
};

struct AllocBufs {
    float* desired;
    unsigned char* clipped;
    AllocState* st;
    SeqSum seq;
    // min and max importance over the unclipped pixels of each chunk of 64
    // pixels and of each super-chunk of 64 chunks (NaN: none left)
    float* cmin;
    float* cmax;
    float* smin;
    float* smax;
};
constexpr int ACH = 64;  // pixels per chunk, chunks per super-chunk

__global__ void k_alloc_begin(int n, AllocConfig ac, int* counts, AllocBufs a) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int floor_pp = min(max(ac.base_samples, 0), ac.adaptive_spp);
    if (i < n) { counts[i] = floor_pp; a.desired[i] = 0.0f; a.clipped[i] = 0; }
    if (i == 0) {
        const int total = ac.adaptive_spp * n;
        const double sum = *a.seq.total;
        const int remaining = max(0, total - floor_pp * n);
        a.st->rem_imp = sum;
        a.st->remaining = remaining;
        a.st->first_clip = 0x7fffffff;
        a.st->final_assign = 0;
        a.st->assigned = 0;
        a.st->active = (total > 0 && sum > 0.0 && remaining > 0) ? 1 : 0;
    }
}

// first index whose share exceeds the cap under the initial state (main.cpp:1391-1398)
__global__ void k_alloc_first_clip(int n, AllocConfig ac, const float* imp, AllocBufs a) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || !a.st->active) return;
    const int max_pp = max(4, 8 * max(1, ac.adaptive_spp));
    const double want = double(imp[i]) / a.st->rem_imp * double(a.st->remaining);
    if (want > double(max_pp)) atomicMin(&a.st->first_clip, i);
}

__device__ __forceinline__ float wave_fmin(float v) {
    for (int o = 32; o > 0; o >>= 1) v = fminf(v, __shfl_xor(v, o));
    return v;
}
__device__ __forceinline__ float wave_fmax(float v) {
    for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o));
    return v;
}

// Chunk and super-chunk min/max of the importance before any clip: one wave
// per chunk (or super-chunk).
__global__ void __launch_bounds__(256) k_alloc_chunk_minmax(int n, const float* imp, AllocBufs a) {
    const int c = (blockIdx.x * blockDim.x + threadIdx.x) >> 6, lane = threadIdx.x & 63;
    const int nch = (n + ACH - 1) / ACH;
    if (c >= nch || !a.st->active) return;
    const int i = c * ACH + lane;
    const float v = i < n ? imp[i] : __builtin_nanf("");
    const float mn = wave_fmin(v), mx = wave_fmax(v);
    if (lane == (c & 63)) { a.cmin[c] = mn; a.cmax[c] = mx; }
}
__global__ void __launch_bounds__(256) k_alloc_super_minmax(int n, AllocBufs a) {
    const int s = (blockIdx.x * blockDim.x + threadIdx.x) >> 6, lane = threadIdx.x & 63;
    const int nch = (n + ACH - 1) / ACH, nsc = (nch + ACH - 1) / ACH;
    if (s >= nsc || !a.st->active) return;
    const int c = s * ACH + lane;
    const float mn = wave_fmin(c < nch ? a.cmin[c] : __builtin_nanf(""));
    const float mx = wave_fmax(c < nch ? a.cmax[c] : __builtin_nanf(""));
    if (lane == (s & 63)) { a.smin[s] = mn; a.smax[s] = mx; }
}

// The clip-and-redistribute sweeps (main.cpp:1387-1412), one wave.
//
// Each sweep visits the unclipped pixels in index order, and clips a pixel
// when double(imp) / rem_imp * remaining exceeds the cap, which changes
// rem_imp and remaining for the pixels after it. For fixed (rem_imp,
// remaining) that expression is monotone in imp (every step is a correctly
// rounded division or multiplication by a fixed operand), so a set of pixels
// holds a clip exactly when its smallest or its largest importance clips. The
// sweep therefore tests whole chunks by their min and max, and whole
// super-chunks likewise, and only walks the pixels of a chunk that holds a
// clip; within such a chunk the 64 pixels stay in registers while their clips
// are taken one after another, in order, each under the state left by the one
// before. Chunk min/max are recomputed when a chunk loses pixels; super-chunk
// min/max are refreshed whenever the sweep descends into one (between
// refreshes they can only bound a superset of its pixels, so a stale value
// costs a look, never a missed clip). Every array element is read and written
// by lane (index & 63) only, so the wave needs no fences between lanes.
__global__ void __launch_bounds__(64) k_alloc_sweeps(int n, AllocConfig ac, const float* imp, AllocBufs a) {
    const int lane = threadIdx.x;
    if (!a.st->active) return;
    const int max_pp = max(4, 8 * max(1, ac.adaptive_spp));
    const float qnan = __builtin_nanf("");
    double rem_imp = a.st->rem_imp;
    int remaining = a.st->remaining;
    int active_count = n;
    const int first = a.st->first_clip;
    const int nch = (n + ACH - 1) / ACH, nsc = (nch + ACH - 1) / ACH;
    bool final_assign = false;
    bool first_sweep = true;
    auto clips = [&](float v) { return double(v) / rem_imp * double(remaining) > double(max_pp); };
    while (active_count > 0 && remaining > 0 && rem_imp > double(RPD_EPS_SMALL)) {
        bool clipped_any = false;
        // under the initial state no pixel before `first` clips
        int pos = first_sweep ? min(first, n) : 0;
        first_sweep = false;
        while (pos < n) {
            // the rest of pos's chunk, pixel by pixel
            const int c = pos / ACH;
            const int i = c * ACH + lane;
            const bool in_range = i < n;
            const float v = in_range ? imp[i] : 0.0f;
            bool cl = in_range ? a.clipped[i] != 0 : true;
            bool changed = false;
            // Clips come in runs (about 27 per chunk at 1920x1080), and each changes the
            // state for the next, so they are resolved a round at a time: every pixel that
            // clips under the round's starting state is taken as clipped, one lane-ordered
            // FP64 chain gives each lane the exact state it would see, and every lane
            // re-tests itself under that state. Up to the first lane whose outcome differs
            // from the guess, every outcome is exact (by induction, each lane saw the true
            // state); that lane's own outcome is exact too, and the next round starts after it.
            while (true) {
                const uint64_t guess = __ballot(!cl && i >= pos && clips(v));
                if (!guess) break;
                double I = rem_imp, Ib = rem_imp;   // running state; state before this lane
                int R = remaining, Rb = remaining;
                for (uint64_t m = guess; m; m &= m - 1) {
                    const int k = __ffsll((long long)m) - 1;
                    I -= double(__int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), k)));
                    R -= max_pp;
                    if (lane > k) { Ib = I; Rb = R; }
                }
                const bool mine = (guess >> lane) & 1ull;
                const bool clip = !cl && i >= pos && double(v) / Ib * double(Rb) > double(max_pp);
                const uint64_t miss = __ballot(clip != mine);
                const int fb = miss ? __ffsll((long long)miss) - 1 : 64;
                // taken: the guessed clips before lane fb, and fb itself if it clips
                const bool take = clip && lane <= fb;
                const uint64_t taken = __ballot(take);
                if (take) {
                    cl = true;
                    a.desired[i] = float(max_pp);
                    a.clipped[i] = 1;
                }
                if (taken) {
                    active_count -= __popcll(taken);
                    clipped_any = true;
                    changed = true;
                }
                if (fb == 64) {          // every guess held: the chain's end is the state
                    rem_imp = I;
                    remaining = R;
                    break;
                }
                const int hi = __builtin_amdgcn_readlane(int(uint64_t(__double_as_longlong(Ib)) >> 32), fb);
                const int lo = __builtin_amdgcn_readlane(int(uint32_t(uint64_t(__double_as_longlong(Ib)))), fb);
                rem_imp = __longlong_as_double((long long)((uint64_t(uint32_t(hi)) << 32) | uint64_t(uint32_t(lo))));
                remaining = __builtin_amdgcn_readlane(Rb, fb);
                if ((taken >> fb) & 1ull) {
                    rem_imp -= double(__int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), fb)));
                    remaining -= max_pp;
                }
                pos = c * ACH + fb + 1;
            }
            if (changed) {
                const float mn = wave_fmin(cl ? qnan : v), mx = wave_fmax(cl ? qnan : v);
                if (lane == (c & 63)) { a.cmin[c] = mn; a.cmax[c] = mx; }
            }
            // the next chunk that holds a clip under the current state
            int next = nch;
            int c1 = c + 1;
            while (c1 < nch) {
                const int sc = c1 / ACH;
                const int j = sc * ACH + lane;
                const float mn = j < nch ? a.cmin[j] : qnan, mx = j < nch ? a.cmax[j] : qnan;
                const float smn = wave_fmin(mn), smx = wave_fmax(mx);
                if (lane == (sc & 63)) { a.smin[sc] = smn; a.smax[sc] = smx; }
                const uint64_t b = __ballot(j < nch && j >= c1 && (clips(mn) || clips(mx)));
                if (b) { next = sc * ACH + __ffsll((long long)b) - 1; break; }
                // later super-chunks, 64 per look, aligned so lane k & 63 owns super-chunk k
                int s1 = sc + 1, found = -1;
                for (int s0 = s1 & ~63; s0 < nsc && found < 0; s0 += 64) {
                    const int k = s0 + lane;
                    const bool ok = k >= s1 && k < nsc;
                    const float kmn = ok ? a.smin[k] : qnan, kmx = ok ? a.smax[k] : qnan;
                    const uint64_t bs = __ballot(ok && (clips(kmn) || clips(kmx)));
                    if (bs) found = s0 + __ffsll((long long)bs) - 1;
                }
                if (found < 0) break;
                c1 = found * ACH;
            }
            if (next >= nch) break;
            pos = next * ACH;
        }
        if (!clipped_any) { final_assign = true; break; }
    }
    if (lane == 0) {
        a.st->rem_imp = rem_imp;
        a.st->remaining = remaining;
        a.st->final_assign = final_assign ? 1 : 0;
    }
}

// shares, floors and residuals (main.cpp:1403-1408, 1414-1423). The floors are
// added to the assigned count as an integer sum, the same in any order: each
// thread sums its pixels (a grid-stride loop over ALLOC_FLOOR_BLOCKS blocks),
// the block sums its threads, and one atomic per block adds that. One atomic
// per wave (16,384 on 1 Mpx, all on one address) made this kernel 190 us.
constexpr int ALLOC_FLOOR_BLOCKS = 512;
__global__ void __launch_bounds__(256) k_alloc_floor(int n, const float* imp, int* counts, AllocBufs a,
                                                     float* resid) {
    __shared__ int s_sum[256 / 64];
    const bool active = a.st->active != 0;
    const bool final_assign = a.st->final_assign != 0;
    const double rem_imp = a.st->rem_imp, remaining = double(a.st->remaining);
    int add = 0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        if (!active) { resid[i] = 0.0f; continue; }
        float d = a.desired[i];
        if (final_assign && !a.clipped[i]) d = float(double(imp[i]) / rem_imp * remaining);
        const int f = int(floorf(d));
        counts[i] += f;
        resid[i] = fmaxf(0.0f, d - float(f));
        add += f;
    }
    for (int o = 32; o > 0; o >>= 1) add += __shfl_xor(add, o);
    if ((threadIdx.x & 63) == 0) s_sum[threadIdx.x >> 6] = add;
    __syncthreads();
    if (threadIdx.x == 0 && active) {
        int sum = 0;
        for (int w = 0; w < 256 / 64; ++w) sum += s_sum[w];
        if (sum) atomicAdd(&a.st->assigned, sum);
    }
}

// systematic residual sampling (main.cpp:1425-1438), counted per pixel
__device__ __forceinline__ double sys_target(int k, double off, int left, double rsum) {
    return ((double(k) + off) / double(left)) * rsum;
}
__device__ __forceinline__ int targets_at_most(double v, double off, int left, double rsum) {
    const double est = v / rsum * double(left) - off;
    int k;
    if (!(est > 0.0)) k = 0;
    else if (est >= double(left)) k = left;
    else k = int(est);
    while (k < left && sys_target(k, off, left, rsum) <= v) ++k;
    while (k > 0 && sys_target(k - 1, off, left, rsum) > v) --k;
    return k;
}
__global__ void k_alloc_systematic(int n, AllocConfig ac, int* counts, AllocBufs a) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n || !a.st->active) return;
    const int total = ac.adaptive_spp * n;
    const int floor_pp = min(max(ac.base_samples, 0), ac.adaptive_spp);
    const int left = max(0, total - floor_pp * n - a.st->assigned);
    const double rsum = *a.seq.total;
    if (left <= 0 || !(rsum > double(RPD_EPS_SMALL))) return;
    const double off = double(hash_to_unit_float(ac.seed ^ 0x4d595df4u));
    const double* P = a.seq.approx;   // exact prefix of the residuals
    const int hi = targets_at_most(P[c], off, left, rsum);
    const int lo = c > 0 ? targets_at_most(P[c - 1], off, left, rsum) : 0;
    if (hi > lo) counts[c] += hi - lo;
}

// ---------------------------------------------------------------------------
static size_t al(size_t b) { return (b + 255) & ~size_t(255); }

size_t alloc_scratch_bytes(int n) {
    const size_t N = size_t(n) + 1;
    const size_t nb = N / SB + 2;
    const size_t nch = N / ACH + 2, nsc = nch / ACH + 2;
    return al(N * 4) + al(N) + al(sizeof(AllocState)) + al(N * 4) + al(N * 4) + al(N * 8) + al(N * 8) +
           al(N * 4) + al(N * 4) + al(N * 4) + al(N * 8) + al(N * 8) + al(8) + al(4) + al(nb * 8) + al(nb * 8) +
           al(nb * 4) + 2 * al(nch * 4) + 2 * al(nsc * 4);
}

struct Carved {
    AllocBufs a;
    float* resid;
    float* tmp;
};
static Carved carve(void* scratch, int n) {
    const size_t N = size_t(n) + 1;
    const size_t nb = N / SB + 2;
    char* p = static_cast<char*>(scratch);
    auto take = [&](size_t bytes) { char* q = p; p += al(bytes); return q; };
    Carved c;
    c.a.desired = reinterpret_cast<float*>(take(N * 4));
    c.a.clipped = reinterpret_cast<unsigned char*>(take(N));
    c.a.st = reinterpret_cast<AllocState*>(take(sizeof(AllocState)));
    c.resid = reinterpret_cast<float*>(take(N * 4));
    c.tmp = reinterpret_cast<float*>(take(N * 4));
    SeqSum& s = c.a.seq;
    s.n = n;
    s.x = nullptr;
    s.approx = reinterpret_cast<double*>(take(N * 8));
    s.units = reinterpret_cast<uint64_t*>(take(N * 8));
    s.serial_pos = reinterpret_cast<int*>(take(N * 4));
    s.serial_list = reinterpret_cast<int*>(take(N * 4));
    s.binade = reinterpret_cast<int*>(take(N * 4));
    s.seg_add = reinterpret_cast<double*>(take(N * 8));
    s.s_at = reinterpret_cast<double*>(take(N * 8));
    s.total = reinterpret_cast<double*>(take(8));
    s.n_serial = reinterpret_cast<int*>(take(4));
    s.bsum_d = reinterpret_cast<double*>(take(nb * 8));
    s.bsum_u = reinterpret_cast<uint64_t*>(take(nb * 8));
    s.bsum_i = reinterpret_cast<int*>(take(nb * 4));
    const size_t nch = N / ACH + 2, nsc = nch / ACH + 2;
    c.a.cmin = reinterpret_cast<float*>(take(nch * 4));
    c.a.cmax = reinterpret_cast<float*>(take(nch * 4));
    c.a.smin = reinterpret_cast<float*>(take(nsc * 4));
    c.a.smax = reinterpret_cast<float*>(take(nsc * 4));
    return c;
}

void launch_importance(int W, int H, const AllocConfig& ac, const PixelBuffers& pb, const Features& f,
                       const float* vden, float* imp, void* scratch, hipStream_t st) {
    const int n = W * H;
    const Carved c = carve(scratch, n);
    const dim3 grid((n + 255) / 256);
    if (ac.adaptive_spp <= 0) return;
    hipLaunchKernelGGL(k_importance, grid, dim3(256), 0, st, n, pb, vden, ac.sigma_max, imp, ac.dbg_variance_source);
    if (ac.radius > 0) {
        const dim3 tiles((W + SMOOTH_TW - 1) / SMOOTH_TW, (H + SMOOTH_TH - 1) / SMOOTH_TH);
        const size_t lds = smooth_lds_bytes(ac.radius);
        hipLaunchKernelGGL(k_importance_smooth, tiles, dim3(SMOOTH_TW * SMOOTH_TH), lds, st, W, H, ac.radius, pb, f,
                           imp, c.tmp, lds > 0 ? 1 : 0);
        hipMemcpyAsync(imp, c.tmp, (size_t)n * sizeof(float), hipMemcpyDeviceToDevice, st);
    }
}

void launch_alloc_counts(int n, const AllocConfig& ac, const float* imp, int* counts, void* scratch, hipStream_t st) {
    Carved c = carve(scratch, n);
    const dim3 g((n + 255) / 256), b(256);
    // sum of the importance field as std::accumulate forms it (main.cpp:1370)
    c.a.seq.x = imp;
    seq_sum_launch(c.a.seq, false, st);
    hipLaunchKernelGGL(k_alloc_begin, g, b, 0, st, n, ac, counts, c.a);
    hipLaunchKernelGGL(k_alloc_first_clip, g, b, 0, st, n, ac, imp, c.a);
    const int nch = (n + ACH - 1) / ACH, nsc = (nch + ACH - 1) / ACH;
    hipLaunchKernelGGL(k_alloc_chunk_minmax, dim3((nch + 3) / 4), dim3(256), 0, st, n, imp, c.a);
    hipLaunchKernelGGL(k_alloc_super_minmax, dim3((nsc + 3) / 4), dim3(256), 0, st, n, c.a);
    hipLaunchKernelGGL(k_alloc_sweeps, dim3(1), dim3(64), 0, st, n, ac, imp, c.a);
    hipLaunchKernelGGL(k_alloc_floor, dim3(std::min<int>((n + 255) / 256, ALLOC_FLOOR_BLOCKS)), b, 0, st, n, imp,
                       counts, c.a, c.resid);
    // residual prefix sums as the reference's running `cumulative` forms them
    c.a.seq.x = c.resid;
    seq_sum_launch(c.a.seq, true, st);
    hipLaunchKernelGGL(k_alloc_systematic, g, b, 0, st, n, ac, counts, c.a);
}

}  // namespace rpd
