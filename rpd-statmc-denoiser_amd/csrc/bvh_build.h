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

// The same layout built on the device (bvh_gpu.hip) from device arrays of the
// primitives, on stream `st`: nodes needs room for n_tris + n_spheres nodes,
// slots and slot_prim one entry per primitive, scratch
// bvh4_device_scratch_bytes(n) bytes. read_back is a synchronous device-to-host
// copy. Returns 0 (with the node count and the 4-wide depth), -1 for an empty
// scene, -2 if the tree is deeper than BVH_MAX_DEPTH4, or a hipError_t code.
size_t bvh4_device_scratch_bytes(int n_prims);
int build_bvh4_device(const DevTriangle* tris, int n_tris, const DevSphere* spheres, int n_spheres, Bvh4Node* nodes,
                      float4* slots, int* slot_prim, void* scratch, int* out_nodes, int* out_depth,
                      hipError_t (*read_back)(void* dst, const void* src, size_t bytes), hipStream_t st);

}  // namespace rpd
