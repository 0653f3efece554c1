// bvh_build.h -- host binned-SAH BVH builder producing the device layout of scene_types.h.
#pragma once

#include <stdint.h>

#include <vector>

namespace sspt {

struct BvhInput {
    // triangle vertex positions, 9 floats per triangle, in global triangle order
    const float* tri_vertices;
    int num_tris;
};

struct BvhOutput {
    std::vector<float> nodes;     // 32 floats per node (8 x float4), 4-wide
    int stack_need = 0;           // largest traversal stack the tree can require
    int depth4 = 0;               // depth of the 4-wide tree
    std::vector<int> leaf_order;  // leaf_order[i] = global triangle id stored at leaf index i
    int depth = 0;
    float scene_lo[3], scene_hi[3];
};

// Triangles per leaf at most (the builder splits every larger range).
#ifndef SSPT_MAX_LEAF
#define SSPT_MAX_LEAF 1
#endif
constexpr int kMaxLeafTriangles = SSPT_MAX_LEAF;

// Binned SAH (SSPT_SAH_BINS bins on centroid bounds), leaves of at most
// `max_leaf` triangles, depth bounded by kMaxBvhDepth. Uses up to `threads`
// host threads.
void build_bvh(const BvhInput& in, int max_leaf, int threads, BvhOutput* out);

}  // namespace sspt
