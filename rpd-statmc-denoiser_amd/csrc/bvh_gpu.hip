// Device-side BVH construction (LBVH): Morton codes of the primitive
// centroids, a radix sort, the binary hierarchy of the sorted codes (Karras,
// "Maximizing parallelism in the construction of BVHs, octrees, and k-d
// trees", HPG 2012), bounding boxes bottom-up, subtrees of at most four
// primitives as leaves, and a collapse into 4-wide nodes level by level. The output has the host builder's
// layout (Bvh4Node, primitive slots in leaf order), so bvh_query traverses it
// unchanged; the host's binned SAH gives better trees, this one is built in
// milliseconds.
#include "bvh_build.h"
#include "kernels.h"

#include <algorithm>
#include <cmath>

namespace rpd {
namespace {

constexpr int LB_BLOCK = 256;
constexpr int LB_LEAF_MAX = 4;  // as the host builder's leaves

__device__ __forceinline__ float lb_fmin(float a, float b) { return a < b ? a : b; }
__device__ __forceinline__ float lb_fmax(float a, float b) { return a > b ? a : b; }

// Primitive boxes (the host builder's: triangle vertex bounds; spheres padded
// so that rounding of centre +- radius never cuts them) and centroids, with a
// per-block min/max of the centroids.
__global__ void __launch_bounds__(LB_BLOCK) k_lb_prims(const DevTriangle* tris, int nt, const DevSphere* sph, int ns,
                                                       float* box, float* cen, float* part) {
    __shared__ float red[6][LB_BLOCK];
    const int i = blockIdx.x * LB_BLOCK + threadIdx.x;
    const int n = nt + ns;
    float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    float c[3] = {0, 0, 0};
    if (i < n) {
        if (i < nt) {
            const float* v[3] = {tris[i].v0, tris[i].v1, tris[i].v2};
            for (int q = 0; q < 3; ++q)
                for (int k = 0; k < 3; ++k) { lo[k] = fminf(lo[k], v[q][k]); hi[k] = fmaxf(hi[k], v[q][k]); }
        } else {
            const DevSphere& s = sph[i - nt];
            for (int k = 0; k < 3; ++k) {
                const float pad = fabsf(s.r) * (1.0f + 1e-5f) + fabsf(s.c[k]) * 1e-6f;
                lo[k] = s.c[k] - pad;
                hi[k] = s.c[k] + pad;
            }
        }
        for (int k = 0; k < 3; ++k) {
            box[6 * size_t(i) + k] = lo[k];
            box[6 * size_t(i) + 3 + k] = hi[k];
            c[k] = 0.5f * (lo[k] + hi[k]);
            cen[3 * size_t(i) + k] = c[k];
        }
    }
    for (int k = 0; k < 3; ++k) {
        red[k][threadIdx.x] = i < n ? c[k] : INFINITY;
        red[3 + k][threadIdx.x] = i < n ? c[k] : -INFINITY;
    }
    __syncthreads();
    for (int o = LB_BLOCK / 2; o > 0; o >>= 1) {
        if (threadIdx.x < o)
            for (int k = 0; k < 3; ++k) {
                red[k][threadIdx.x] = lb_fmin(red[k][threadIdx.x], red[k][threadIdx.x + o]);
                red[3 + k][threadIdx.x] = lb_fmax(red[3 + k][threadIdx.x], red[3 + k][threadIdx.x + o]);
            }
        __syncthreads();
    }
    if (threadIdx.x < 6) part[6 * blockIdx.x + threadIdx.x] = red[threadIdx.x][0];
}

// the centroid bounds of the scene from the per-block parts (one block)
__global__ void __launch_bounds__(LB_BLOCK) k_lb_bounds(int nparts, const float* part, float* scene) {
    __shared__ float red[6][LB_BLOCK];
    float v[6] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
    for (int b = threadIdx.x; b < nparts; b += LB_BLOCK)
        for (int k = 0; k < 3; ++k) {
            v[k] = lb_fmin(v[k], part[6 * b + k]);
            v[3 + k] = lb_fmax(v[3 + k], part[6 * b + 3 + k]);
        }
    for (int k = 0; k < 6; ++k) red[k][threadIdx.x] = v[k];
    __syncthreads();
    for (int o = LB_BLOCK / 2; o > 0; o >>= 1) {
        if (threadIdx.x < o)
            for (int k = 0; k < 3; ++k) {
                red[k][threadIdx.x] = lb_fmin(red[k][threadIdx.x], red[k][threadIdx.x + o]);
                red[3 + k][threadIdx.x] = lb_fmax(red[3 + k][threadIdx.x], red[3 + k][threadIdx.x + o]);
            }
        __syncthreads();
    }
    if (threadIdx.x < 6) scene[threadIdx.x] = red[threadIdx.x][0];
}

// 10 bits per axis, interleaved (x highest)
__device__ __forceinline__ uint32_t lb_spread10(uint32_t v) {
    v &= 0x3ffu;
    v = (v | (v << 16)) & 0x030000ffu;
    v = (v | (v << 8)) & 0x0300f00fu;
    v = (v | (v << 4)) & 0x030c30c3u;
    v = (v | (v << 2)) & 0x09249249u;
    return v;
}
__global__ void __launch_bounds__(LB_BLOCK) k_lb_morton(int n, const float* cen, const float* scene, uint32_t* keys,
                                                        int* vals) {
    const int i = blockIdx.x * LB_BLOCK + threadIdx.x;
    if (i >= n) return;
    uint32_t q[3];
    for (int k = 0; k < 3; ++k) {
        const float ext = scene[3 + k] - scene[k];
        const float u = ext > 0.0f ? (cen[3 * size_t(i) + k] - scene[k]) / ext : 0.5f;
        q[k] = uint32_t(fminf(fmaxf(u * 1024.0f, 0.0f), 1023.0f));
    }
    keys[i] = (lb_spread10(q[0]) << 2) | (lb_spread10(q[1]) << 1) | lb_spread10(q[2]);
    vals[i] = i;
}

// Length of the common prefix of sorted entries i and j, with the position as
// a tie-break so that equal codes still split (-1 outside the range).
__device__ __forceinline__ int lb_delta(const uint32_t* keys, int n, int i, int j) {
    if (j < 0 || j >= n) return -1;
    const uint32_t a = keys[i], b = keys[j];
    if (a == b) return 32 + __clz(uint32_t(i) ^ uint32_t(j));
    return __clz(a ^ b);
}

// Internal node i of the n - 1: its range [first, first + cnt) of sorted
// entries and its two children (>= 0: internal node, < 0: ~sorted entry).
__global__ void __launch_bounds__(LB_BLOCK) k_lb_karras(int n, const uint32_t* keys, int* left, int* right,
                                                        int* first, int* cnt, int* parent, int* leaf_parent) {
    const int i = blockIdx.x * LB_BLOCK + threadIdx.x;
    if (i >= n - 1) return;
    const int d = lb_delta(keys, n, i, i + 1) - lb_delta(keys, n, i, i - 1) >= 0 ? 1 : -1;
    const int dmin = lb_delta(keys, n, i, i - d);
    int lmax = 2;
    while (lb_delta(keys, n, i, i + lmax * d) > dmin) lmax *= 2;
    int l = 0;
    for (int t = lmax / 2; t >= 1; t /= 2)
        if (lb_delta(keys, n, i, i + (l + t) * d) > dmin) l += t;
    const int j = i + l * d;
    const int dnode = lb_delta(keys, n, i, j);
    int s = 0;
    for (int div = 2;; div *= 2) {
        const int t = (l + div - 1) / div;
        if (lb_delta(keys, n, i, i + (s + t) * d) > dnode) s += t;
        if (t <= 1) break;
    }
    const int g = i + s * d + (d < 0 ? -1 : 0);
    const int lo = min(i, j), hi = max(i, j);
    const int a = (lo == g) ? ~g : g;
    const int b = (hi == g + 1) ? ~(g + 1) : g + 1;
    left[i] = a;
    right[i] = b;
    first[i] = lo;
    cnt[i] = hi - lo + 1;
    if (a >= 0) parent[a] = i; else leaf_parent[~a] = i;
    if (b >= 0) parent[b] = i; else leaf_parent[~b] = i;
    if (i == 0) parent[0] = -1;
}

__device__ __forceinline__ void lb_child_box(int c, const int* vals, const float* pbox, const float* nbox,
                                             float out[6]) {
    const float* src = c < 0 ? pbox + 6 * size_t(vals[~c]) : nbox + 6 * size_t(c);
    for (int k = 0; k < 6; ++k) out[k] = src[k];
}

// Boxes bottom-up: one thread per sorted entry climbs while it is the second
// child to arrive at a node (the first one stops), so every box is formed from
// two finished children.
__global__ void __launch_bounds__(LB_BLOCK) k_lb_refit(int n, const int* vals, const float* pbox, const int* left,
                                                       const int* right, const int* parent, const int* leaf_parent,
                                                       float* nbox, int* flags) {
    const int s = blockIdx.x * LB_BLOCK + threadIdx.x;
    if (s >= n) return;
    int p = leaf_parent[s];
    while (p >= 0) {
        __threadfence();
        if (atomicAdd(flags + p, 1) == 0) return;
        __threadfence();
        float a[6], b[6];
        lb_child_box(left[p], vals, pbox, nbox, a);
        lb_child_box(right[p], vals, pbox, nbox, b);
        for (int k = 0; k < 3; ++k) {
            nbox[6 * size_t(p) + k] = fminf(a[k], b[k]);
            nbox[6 * size_t(p) + 3 + k] = fmaxf(a[3 + k], b[3 + k]);
        }
        p = parent[p];
    }
}

__device__ __forceinline__ void lb_set_child(Bvh4Node& nd, int k, const float b[6], int child, int count) {
    nd.lo_x[k] = b[0]; nd.hi_x[k] = b[3];
    nd.lo_y[k] = b[1]; nd.hi_y[k] = b[4];
    nd.lo_z[k] = b[2]; nd.hi_z[k] = b[5];
    nd.child[k] = child;
    nd.count[k] = count;
}

struct LbTree {
    const int* vals;
    const float* pbox;
    const float* nbox;
    const int* left;
    const int* right;
    const int* first;
    const int* cnt;
};

// a child (binary) that becomes a leaf of the 4-wide tree: one entry, or a subtree of at most four
__device__ __forceinline__ bool lb_is_leaf(const LbTree& t, int c) { return c < 0 || t.cnt[c] <= LB_LEAF_MAX; }

// One level of the collapse: every frontier entry (binary node, 4-wide index)
// becomes a 4-wide node over up to four of its descendants; the inner ones
// get new indices and form the next frontier. Counters: [0] nodes emitted, [1 + level % 2] frontier sizes (the
// next level's is cleared by k_lb_clear before the launch); lvl[level]
// records this level's frontier size.
__global__ void __launch_bounds__(LB_BLOCK) k_lb_collapse(LbTree t, const int2* front_in, int2* front_out,
                                                          int* ctr, int* lvl, int level, Bvh4Node* nodes) {
    const int n_in = ctr[1 + (level & 1)];
    if (blockIdx.x == 0 && threadIdx.x == 0) lvl[level] = n_in;
    for (int k = blockIdx.x * LB_BLOCK + threadIdx.x; k < n_in; k += gridDim.x * LB_BLOCK) {
        const int b = front_in[k].x, o = front_in[k].y;
        // children: start from the two of b and split the one with the most
        // primitives until there are four (Morton splits are often lopsided,
        // so this keeps the 4-wide tree shallower than taking grandchildren)
        int kids[4] = {t.left[b], t.right[b], 0, 0};
        int nk = 2;
        while (nk < 4) {
            int pick = -1, most = LB_LEAF_MAX;
            for (int q = 0; q < nk; ++q)
                if (!lb_is_leaf(t, kids[q]) && t.cnt[kids[q]] > most) { most = t.cnt[kids[q]]; pick = q; }
            if (pick < 0) break;
            const int c = kids[pick];
            kids[pick] = t.left[c];
            kids[nk++] = t.right[c];
        }
        Bvh4Node nd;
        const float empty[6] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
        for (int q = 0; q < 4; ++q) lb_set_child(nd, q, empty, -1, -1);
        for (int q = 0; q < nk; ++q) {
            const int c = kids[q];
            float bx[6];
            lb_child_box(c, t.vals, t.pbox, t.nbox, bx);
            if (lb_is_leaf(t, c)) {
                lb_set_child(nd, q, bx, c < 0 ? ~c : t.first[c], c < 0 ? 1 : t.cnt[c]);
            } else {
                const int idx = atomicAdd(ctr, 1);
                const int at = atomicAdd(ctr + 1 + ((level + 1) & 1), 1);
                front_out[at] = make_int2(c, idx);
                lb_set_child(nd, q, bx, idx, 0);
            }
        }
        nodes[o] = nd;
    }
}

// primitive slots in sorted (leaf) order, as the host builder lays them out
__global__ void __launch_bounds__(LB_BLOCK) k_lb_slots(int n, int nt, const int* vals, const DevTriangle* tris,
                                                       const DevSphere* sph, float4* slots, int* slot_prim) {
    const int s = blockIdx.x * LB_BLOCK + threadIdx.x;
    if (s >= n) return;
    const int p = vals[s];
    float4* e = slots + size_t(s) * BVH_SLOT_FLOAT4;
    if (p < nt) {
        slot_prim[s] = p;
        const float* v[3] = {tris[p].v0, tris[p].v1, tris[p].v2};
        for (int kz = 0; kz < 3; ++kz) {
            const int kx = kz == 2 ? 0 : kz + 1, ky = kx == 2 ? 0 : kx + 1;
            float f[12] = {};
            for (int q = 0; q < 3; ++q) { f[3 * q] = v[q][kx]; f[3 * q + 1] = v[q][ky]; f[3 * q + 2] = v[q][kz]; }
            for (int q = 0; q < 3; ++q) e[kz * 3 + q] = make_float4(f[4 * q], f[4 * q + 1], f[4 * q + 2], f[4 * q + 3]);
        }
    } else {
        const DevSphere& sp = sph[p - nt];
        slot_prim[s] = (p - nt) | int(0x80000000u);
        e[0] = make_float4(sp.c[0], sp.c[1], sp.c[2], sp.r);
        for (int q = 1; q < BVH_SLOT_FLOAT4; ++q) e[q] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    }
}

// a whole scene that fits one leaf: one node with one leaf child
__global__ void k_lb_single(int n, const float* pbox, const int* vals, Bvh4Node* nodes) {
    float b[6] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
    for (int s = 0; s < n; ++s)
        for (int k = 0; k < 3; ++k) {
            b[k] = fminf(b[k], pbox[6 * size_t(vals[s]) + k]);
            b[3 + k] = fmaxf(b[3 + k], pbox[6 * size_t(vals[s]) + 3 + k]);
        }
    Bvh4Node nd;
    const float empty[6] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
    for (int q = 0; q < 4; ++q) lb_set_child(nd, q, empty, -1, -1);
    lb_set_child(nd, 0, b, 0, n);
    nodes[0] = nd;
}

__global__ void k_lb_clear(int* p) { *p = 0; }

__global__ void k_lb_root(int2* front, int* ctr) {
    front[0] = make_int2(0, 0);
    ctr[0] = 1;   // node 0 is the root
    ctr[1] = 1;   // level 0 frontier: the root
    ctr[2] = 0;
}

__global__ void k_lb_fill_i32(int* p, int v, int n) {
    const int i = blockIdx.x * LB_BLOCK + threadIdx.x;
    if (i < n) p[i] = v;
}

size_t lb_al(size_t b) { return (b + 255) / 256 * 256; }

struct LbScratch {
    float* pbox; float* cen; float* part; float* scene;
    uint32_t* keys; uint32_t* keys_s; int* vals; int* vals_s;
    int *left, *right, *first, *cnt, *parent, *leaf_parent, *flags;
    float* nbox;
    int2* front[2];
    int* ctr;
    int* lvl;   // BVH_MAX_DEPTH4 frontier sizes
    void* sort_tmp; size_t sort_bytes;
};

LbScratch lb_carve(void* mem, int n) {
    const size_t N = size_t(std::max(n, 1));
    const size_t nparts = (N + LB_BLOCK - 1) / LB_BLOCK;
    char* c = static_cast<char*>(mem);
    auto take = [&](size_t b) { char* r = c; c += lb_al(b); return r; };
    LbScratch s;
    s.pbox = reinterpret_cast<float*>(take(N * 24));
    s.cen = reinterpret_cast<float*>(take(N * 12));
    s.part = reinterpret_cast<float*>(take(nparts * 24));
    s.scene = reinterpret_cast<float*>(take(24));
    s.keys = reinterpret_cast<uint32_t*>(take(N * 4));
    s.keys_s = reinterpret_cast<uint32_t*>(take(N * 4));
    s.vals = reinterpret_cast<int*>(take(N * 4));
    s.vals_s = reinterpret_cast<int*>(take(N * 4));
    s.left = reinterpret_cast<int*>(take(N * 4));
    s.right = reinterpret_cast<int*>(take(N * 4));
    s.first = reinterpret_cast<int*>(take(N * 4));
    s.cnt = reinterpret_cast<int*>(take(N * 4));
    s.parent = reinterpret_cast<int*>(take(N * 4));
    s.leaf_parent = reinterpret_cast<int*>(take(N * 4));
    s.flags = reinterpret_cast<int*>(take(N * 4));
    s.nbox = reinterpret_cast<float*>(take(N * 24));
    s.front[0] = reinterpret_cast<int2*>(take(N * 8));
    s.front[1] = reinterpret_cast<int2*>(take(N * 8));
    s.ctr = reinterpret_cast<int*>(take(16));
    s.lvl = reinterpret_cast<int*>(take(BVH_MAX_DEPTH4 * 4));
    s.sort_bytes = sort_pairs_u32_temp_bytes(int(N));
    s.sort_tmp = take(std::max<size_t>(s.sort_bytes, 16));
    return s;
}

}  // namespace

size_t bvh4_device_scratch_bytes(int n) {
    const size_t N = size_t(std::max(n, 1));
    const size_t nparts = (N + LB_BLOCK - 1) / LB_BLOCK;
    return lb_al(N * 24) + lb_al(N * 12) + lb_al(nparts * 24) + lb_al(24) + 4 * lb_al(N * 4) + 7 * lb_al(N * 4) +
           lb_al(N * 24) + 2 * lb_al(N * 8) + lb_al(16) + lb_al(BVH_MAX_DEPTH4 * 4) + lb_al(std::max<size_t>(sort_pairs_u32_temp_bytes(int(N)), 16));
}

int build_bvh4_device(const DevTriangle* tris, int n_tris, const DevSphere* spheres, int n_spheres, Bvh4Node* nodes,
                      float4* slots, int* slot_prim, void* scratch, int* out_nodes, int* out_depth,
                      hipError_t (*read_back)(void* dst, const void* src, size_t bytes), hipStream_t st) {
    const int n = n_tris + n_spheres;
    if (n <= 0) return -1;
    LbScratch s = lb_carve(scratch, n);
    const int g = (n + LB_BLOCK - 1) / LB_BLOCK;
    hipLaunchKernelGGL(k_lb_prims, dim3(g), dim3(LB_BLOCK), 0, st, tris, n_tris, spheres, n_spheres, s.pbox, s.cen,
                       s.part);
    hipLaunchKernelGGL(k_lb_bounds, dim3(1), dim3(LB_BLOCK), 0, st, g, s.part, s.scene);
    hipLaunchKernelGGL(k_lb_morton, dim3(g), dim3(LB_BLOCK), 0, st, n, s.cen, s.scene, s.keys, s.vals);
    hipError_t e = sort_pairs_u32(s.keys, s.keys_s, s.vals, s.vals_s, n, 30, s.sort_tmp, s.sort_bytes, st);
    if (e != hipSuccess) return int(e);
    hipLaunchKernelGGL(k_lb_slots, dim3(g), dim3(LB_BLOCK), 0, st, n, n_tris, s.vals_s, tris, spheres, slots,
                       slot_prim);
    int depth = 1, count = 1;
    if (n <= LB_LEAF_MAX) {
        hipLaunchKernelGGL(k_lb_single, dim3(1), dim3(1), 0, st, n, s.pbox, s.vals_s, nodes);
    } else {
        hipLaunchKernelGGL(k_lb_karras, dim3(g), dim3(LB_BLOCK), 0, st, n, s.keys_s, s.left, s.right, s.first, s.cnt,
                           s.parent, s.leaf_parent);
        hipLaunchKernelGGL(k_lb_fill_i32, dim3(g), dim3(LB_BLOCK), 0, st, s.flags, 0, n);
        hipLaunchKernelGGL(k_lb_refit, dim3(g), dim3(LB_BLOCK), 0, st, n, s.vals_s, s.pbox, s.left, s.right,
                           s.parent, s.leaf_parent, s.nbox, s.flags);
        hipLaunchKernelGGL(k_lb_root, dim3(1), dim3(1), 0, st, s.front[0], s.ctr);
        const LbTree t{s.vals_s, s.pbox, s.nbox, s.left, s.right, s.first, s.cnt};
        const int cg = std::min(g, 1024);
        for (int level = 0; level < BVH_MAX_DEPTH4; ++level) {
            hipLaunchKernelGGL(k_lb_clear, dim3(1), dim3(1), 0, st, s.ctr + 1 + ((level + 1) & 1));
            hipLaunchKernelGGL(k_lb_collapse, dim3(cg), dim3(LB_BLOCK), 0, st, t, s.front[level & 1],
                               s.front[(level + 1) & 1], s.ctr, s.lvl, level, nodes);
        }
        int c[3] = {0, 0, 0}, lv[BVH_MAX_DEPTH4] = {};
        if ((e = hipStreamSynchronize(st)) != hipSuccess) return int(e);
        if ((e = read_back(c, s.ctr, sizeof(c))) != hipSuccess) return int(e);
        if ((e = read_back(lv, s.lvl, sizeof(lv))) != hipSuccess) return int(e);
        count = c[0];
        // a frontier still waiting after BVH_MAX_DEPTH4 levels: too deep for the traversal stack
        if (c[1 + (BVH_MAX_DEPTH4 & 1)] != 0) return -2;
        depth = 0;
        while (depth < BVH_MAX_DEPTH4 && lv[depth] > 0) ++depth;
    }
    if ((e = hipStreamSynchronize(st)) != hipSuccess) return int(e);
    *out_nodes = count;
    *out_depth = depth;
    return 0;
}

}  // namespace rpd
