// scene_types.h -- device-resident scene layout (HBM), shared by host and device code.
//
// Triangles are stored once, in BVH leaf order, so the index the traversal
// returns addresses the shading data directly:
//   tri_pos[3*i+k]  float4  vertex k position; tri_pos[3*i].w = global triangle id (bits)
//   tri_aux[4*i+k]  float4  k<3: shading normal of vertex k (xyz), w = uv component
//                           (w0 = u0, w1 = v0, w2 = u1); tri_aux[4*i+3] = (v1, u2, v2, info bits)
//   info bits: 0..23 material id, bit 24 has normals, bit 25 has uvs
//   tri_emitter[i]  int     emitter index in Scene.Emitters order, -1 if none
// BVH: binary tree whose nodes hold both child boxes (64 B per node):
//   node[4*n+0] = (L.lo.x, L.hi.x, L.lo.y, L.hi.y)
//   node[4*n+1] = (R.lo.x, R.hi.x, R.lo.y, R.hi.y)
//   node[4*n+2] = (L.lo.z, L.hi.z, R.lo.z, R.hi.z)
//   node[4*n+3] = (left, right, left_count, right_count) as int bits;
//                 count > 0: the child is a leaf of `count` triangles starting at index `left`/`right`.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace sspt {

constexpr int kInfoMaterialMask = 0xFFFFFF;
constexpr int kInfoHasNormals = 1 << 24;
constexpr int kInfoHasUvs = 1 << 25;

// BVH depth bound = traversal stack entries per lane (LDS).
constexpr int kMaxBvhDepth = 40;
constexpr int kTraceBlock = 128;

struct DevMaterial {  // 64 B
    int kind;
    float base[3];
    int base_tex;
    float roughness;
    int rough_tex;
    float metallic, spec_tint, anisotropic, spec_trans, ior;
    int transmitter;
    int pad[3];
};

struct DevTexture {
    uint64_t offset;  // float offset into the texel pool
    int w, h, ch, border;
};

struct DevEmitter {  // 32 B
    int tri;  // leaf-order triangle index
    float r, g, b;
    float inv_area;
    int pad[3];
};

struct DevCamera {
    float v2c[16];  // view -> camera (Matrix4x4 row-major, row vectors)
    float c2w[16];  // camera -> world
    float pos[3];
    int width, height;
};

struct DevScene {
    const float4* nodes;
    const float4* tri_pos;
    const float4* tri_aux;
    const int* tri_emitter;
    const DevMaterial* materials;
    const DevTexture* textures;
    const float* texels;
    const DevEmitter* emitters;
    int num_emitters;
    int num_tris;
    float abs_max;  // largest vertex coordinate magnitude (ray slack scale)
    int has_env;
    int env_w, env_h;
    const float4* env_rgb;     // env_w * env_h, w unused
    const float* env_row_cdf;  // env_h
    const float* env_col_cdf;  // env_h * env_w
    DevCamera cam;
    int width, height;
};

// Path-tracer settings as seen by the kernels.
struct DevSettings {
    uint32_t base_seed;
    int max_depth, min_depth, num_shadow_rays, enable_bsdf_di;
};

// Device-side counters. Index names.
enum Counter {
    kCntNext = 0,    // next extension queue length
    kCntShadow = 1,  // shadow queue length
    kCntNum = 2
};
enum Stat { kStatRays = 0, kStatShadow = 1, kStatOccluded = 2, kStatHits = 3, kStatNan = 4, kStatNum = 5 };

}  // namespace sspt
