// Host-side BVH construction for scenes traced through the BVH: a binary
// BVH by binned surface-area heuristic over triangles and spheres, collapsed
// into 4-wide nodes (bvh_query in rpd_device.h traverses them). Primitive
// slots are laid out in leaf order with each triangle's vertices in the three
// axis orders of the watertight test, so a leaf is one contiguous read.
#include "bvh_build.h"

#include <algorithm>
#include <cmath>
#include <cstring>

namespace rpd {
namespace {

struct Box {
    float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    void grow(const Box& b) {
        for (int k = 0; k < 3; ++k) { lo[k] = std::min(lo[k], b.lo[k]); hi[k] = std::max(hi[k], b.hi[k]); }
    }
    void grow(const float* p) {
        for (int k = 0; k < 3; ++k) { lo[k] = std::min(lo[k], p[k]); hi[k] = std::max(hi[k], p[k]); }
    }
    float area() const {
        const float dx = hi[0] - lo[0], dy = hi[1] - lo[1], dz = hi[2] - lo[2];
        if (!(dx >= 0.0f) || !(dy >= 0.0f) || !(dz >= 0.0f)) return 0.0f;
        return 2.0f * (dx * dy + dy * dz + dz * dx);
    }
};

struct BNode {       // binary node
    Box box;
    int left = -1, right = -1;
    int first = 0, count = 0;  // leaf range in the primitive order
    bool leaf() const { return left < 0; }
};

constexpr int kBins = 16;
constexpr int kLeafMax = 4;
constexpr int kMaxBinaryDepth = 2 * BVH_MAX_DEPTH4 - 2;  // 4-wide depth stays within BVH_MAX_DEPTH4

struct Builder {
    std::vector<Box> pbox;
    std::vector<float> cen;   // 3 per primitive
    std::vector<int> order;   // primitive ids (triangles, then spheres as n_tris + k)
    std::vector<BNode> nodes;

    int build(int first, int count, int depth) {
        const int id = int(nodes.size());
        nodes.emplace_back();
        Box box, cbox;
        for (int i = first; i < first + count; ++i) {
            box.grow(pbox[order[i]]);
            cbox.grow(&cen[3 * order[i]]);
        }
        nodes[id].box = box;
        if (count <= kLeafMax || depth >= kMaxBinaryDepth) {
            nodes[id].first = first; nodes[id].count = count;
            return id;
        }
        // best binned SAH split over the three axes
        float best_cost = INFINITY;
        int best_axis = -1, best_split = -1;
        for (int ax = 0; ax < 3; ++ax) {
            const float lo = cbox.lo[ax], ext = cbox.hi[ax] - lo;
            if (!(ext > 0.0f)) continue;
            Box bb[kBins];
            int bc[kBins] = {};
            const float scale = float(kBins) / ext;
            for (int i = first; i < first + count; ++i) {
                const int p = order[i];
                const int b = std::min(kBins - 1, int((cen[3 * p + ax] - lo) * scale));
                bb[b].grow(pbox[p]);
                ++bc[b];
            }
            float ra[kBins];
            int rc[kBins];
            Box acc;
            int n = 0;
            for (int b = kBins - 1; b > 0; --b) { acc.grow(bb[b]); n += bc[b]; ra[b] = acc.area(); rc[b] = n; }
            Box lacc;
            int ln = 0;
            for (int b = 0; b < kBins - 1; ++b) {
                lacc.grow(bb[b]);
                ln += bc[b];
                if (ln == 0 || rc[b + 1] == 0) continue;
                const float cost = float(ln) * lacc.area() + float(rc[b + 1]) * ra[b + 1];
                if (cost < best_cost) { best_cost = cost; best_axis = ax; best_split = b; }
            }
        }
        int mid;
        if (best_axis >= 0) {
            const float lo = cbox.lo[best_axis], scale = float(kBins) / (cbox.hi[best_axis] - lo);
            auto it = std::partition(order.begin() + first, order.begin() + first + count, [&](int p) {
                return std::min(kBins - 1, int((cen[3 * p + best_axis] - lo) * scale)) <= best_split;
            });
            mid = int(it - order.begin());
            if (mid == first || mid == first + count) mid = first + count / 2;
        } else {
            mid = first + count / 2;  // all centroids coincide: split the range
        }
        const int l = build(first, mid - first, depth + 1);
        const int r = build(mid, first + count - mid, depth + 1);
        nodes[id].left = l;
        nodes[id].right = r;
        return id;
    }
};

void set_child(Bvh4Node& n, int k, const Box& b, int child, int count) {
    n.lo_x[k] = b.lo[0]; n.hi_x[k] = b.hi[0];
    n.lo_y[k] = b.lo[1]; n.hi_y[k] = b.hi[1];
    n.lo_z[k] = b.lo[2]; n.hi_z[k] = b.hi[2];
    n.child[k] = child;
    n.count[k] = count;
}

struct Collapser {
    const std::vector<BNode>& bn;
    std::vector<Bvh4Node>& out;
    int max_depth = 0;

    // 4-wide node for binary node b (inner), returns its index
    int emit(int b, int depth) {
        max_depth = std::max(max_depth, depth);
        const int id = int(out.size());
        out.emplace_back();
        // the grandchildren (a leaf child stays itself): every 4-wide level
        // covers two binary levels, so 4-wide depth <= (binary depth + 1) / 2
        std::vector<int> kids;
        for (const int c : {bn[b].left, bn[b].right}) {
            if (bn[c].leaf()) kids.push_back(c);
            else { kids.push_back(bn[c].left); kids.push_back(bn[c].right); }
        }
        Bvh4Node n;
        for (int k = 0; k < 4; ++k) set_child(n, k, Box(), -1, -1);
        for (int k = 0; k < int(kids.size()); ++k) {
            const BNode& c = bn[kids[k]];
            if (c.leaf()) set_child(n, k, c.box, c.first, c.count);
            else set_child(n, k, c.box, 0, 0);
        }
        out[id] = n;
        for (int k = 0; k < int(kids.size()); ++k)
            if (!bn[kids[k]].leaf()) out[id].child[k] = emit(kids[k], depth + 1);
        return id;
    }
};

}  // namespace

int build_bvh4(const DevTriangle* tris, int n_tris, const DevSphere* spheres, int n_spheres, BvhHost& out) {
    const int n = n_tris + n_spheres;
    if (n <= 0) return -1;
    Builder b;
    b.pbox.resize(n);
    b.cen.resize(3 * size_t(n));
    b.order.resize(n);
    for (int i = 0; i < n_tris; ++i) {
        Box bx;
        bx.grow(tris[i].v0); bx.grow(tris[i].v1); bx.grow(tris[i].v2);
        b.pbox[i] = bx;
    }
    for (int i = 0; i < n_spheres; ++i) {
        Box bx;
        const DevSphere& s = spheres[i];
        for (int k = 0; k < 3; ++k) {
            // widened so that float rounding of centre +- radius never cuts the sphere
            const float pad = std::fabs(s.r) * (1.0f + 1e-5f) + std::fabs(s.c[k]) * 1e-6f;
            bx.lo[k] = s.c[k] - pad;
            bx.hi[k] = s.c[k] + pad;
        }
        b.pbox[n_tris + i] = bx;
    }
    for (int i = 0; i < n; ++i) {
        b.order[i] = i;
        for (int k = 0; k < 3; ++k) b.cen[3 * size_t(i) + k] = 0.5f * (b.pbox[i].lo[k] + b.pbox[i].hi[k]);
    }
    b.nodes.reserve(2 * size_t(n) / kLeafMax + 16);
    const int root = b.build(0, n, 0);
    out.nodes.clear();
    Collapser c{b.nodes, out.nodes};
    if (b.nodes[root].leaf()) {  // a handful of primitives: one node with one leaf child
        Bvh4Node nd;
        for (int k = 0; k < 4; ++k) set_child(nd, k, Box(), -1, -1);
        set_child(nd, 0, b.nodes[root].box, b.nodes[root].first, b.nodes[root].count);
        out.nodes.push_back(nd);
        c.max_depth = 1;
    } else {
        c.emit(root, 1);
    }
    out.depth = c.max_depth;
    // primitive slots in leaf order
    out.slots.assign(size_t(n) * BVH_SLOT_FLOAT4, make_float4(0.0f, 0.0f, 0.0f, 0.0f));
    out.slot_prim.resize(n);
    for (int s = 0; s < n; ++s) {
        const int p = b.order[s];
        float4* e = &out.slots[size_t(s) * BVH_SLOT_FLOAT4];
        if (p < n_tris) {
            out.slot_prim[s] = p;
            const float* v[3] = {tris[p].v0, tris[p].v1, tris[p].v2};
            for (int kz = 0; kz < 3; ++kz) {
                const int kx = kz == 2 ? 0 : kz + 1, ky = kx == 2 ? 0 : kx + 1;
                float f[12] = {};
                for (int q = 0; q < 3; ++q) { f[3 * q] = v[q][kx]; f[3 * q + 1] = v[q][ky]; f[3 * q + 2] = v[q][kz]; }
                for (int q = 0; q < 3; ++q) e[kz * 3 + q] = make_float4(f[4 * q], f[4 * q + 1], f[4 * q + 2], f[4 * q + 3]);
            }
        } else {
            const DevSphere& sp = spheres[p - n_tris];
            out.slot_prim[s] = (p - n_tris) | int(0x80000000u);
            e[0] = make_float4(sp.c[0], sp.c[1], sp.c[2], sp.r);
        }
    }
    return out.depth <= BVH_MAX_DEPTH4 ? 0 : -2;
}

}  // namespace rpd
