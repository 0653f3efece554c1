// bvh_build.h -- host binned-SAH BVH builder producing the device layout of scene_types.h.
#pragma once

#include <stdint.h>

#include <string>
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

// SAH bins per axis and split (both builders).
#ifndef SSPT_SAH_BINS
#define SSPT_SAH_BINS 64
#endif
constexpr int kSahBins = SSPT_SAH_BINS;

// Triangles per leaf at most by default (the builder splits every larger range;
// sspt_set_bvh_leaf_size changes it per context).
#ifndef SSPT_MAX_LEAF
#define SSPT_MAX_LEAF 2
#endif
constexpr int kMaxLeafTriangles = SSPT_MAX_LEAF;

// Binned SAH (SSPT_SAH_BINS bins on centroid bounds), leaves of at most
// `max_leaf` triangles, depth bounded by kMaxBvhDepth. Uses up to `threads`
// host threads.
void build_bvh(const BvhInput& in, int max_leaf, int threads, BvhOutput* out);

// The same build on the device (bvh_gpu.hip): the same bins, split costs (fp32,
// same operation order), first-minimum split choice and stable partitions
// (every node's range keeps ascending triangle ids, which fixes the object-median
// split of coincident centroids), hence the same tree, leaf order included; only
// a bound may be -0 where the host has +0. Runs on `stream`; returns false and
// fills `error` on a HIP failure. `device_ms` receives the summed duration of
// the build kernels.
bool build_bvh_device(void* stream, const BvhInput& in, int max_leaf, BvhOutput* out, double* device_ms,
                      std::string* error);

// Traversal nodes (scene_types.h): re-encodes a 4-wide float tree (32 floats
// per node, the builders' layout) as kTravNodeWords words per node, same node
// indices. Leaves of more than four triangles get numbered (first, count)
// pairs in `big_leaves`. Fails (filling `error`) if a leaf starts at a
// triangle index of 2^30 or more.
bool encode_traversal_nodes(const std::vector<float>& nodes, int threads, std::vector<uint32_t>* words,
                            std::vector<int>* big_leaves, std::string* error);

}  // namespace sspt
