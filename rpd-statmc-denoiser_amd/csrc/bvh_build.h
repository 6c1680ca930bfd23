// Host-side BVH construction (bvh_build.hip).
#pragma once
#include <vector>

#include "rpd_device.h"

namespace rpd {

struct BvhHost {
    std::vector<Bvh4Node> nodes;   // node 0 is the root
    std::vector<float4> slots;     // BVH_SLOT_FLOAT4 per primitive slot
    std::vector<int> slot_prim;
    int depth = 0;                 // 4-wide levels
};

// 0 on success; -1 for an empty scene; -2 if the tree got deeper than BVH_MAX_DEPTH4.
int build_bvh4(const DevTriangle* tris, int n_tris, const DevSphere* spheres, int n_spheres, BvhHost& out);

}  // namespace rpd
