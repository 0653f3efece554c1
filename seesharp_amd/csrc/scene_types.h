// scene_types.h -- device-resident scene layout (HBM), shared by host and device code.
//
// One device "world" holds one or more scenes (a batch); all scenes share
// pooled arrays, and a per-scene record gives each its root node, emitter
// range, environment map, camera, settings and framebuffer.
//
// Triangles are stored once, in BVH leaf order, so the index the traversal
// returns addresses the shading data directly:
//   tri_pos[3*i+k]  float4  vertex k position; tri_pos[3*i].w = triangle id within the
//                           world (scene offset + mesh offset + PrimId), int bits
//   tri_aux[4*i+k]  float4  k<3: shading normal of vertex k (xyz), w = uv component
//                           (w0 = u0, w1 = v0, w2 = u1); tri_aux[4*i+3] = (v1, u2, v2, info bits)
//   info bits: 0..23 material id (world-wide), bit 24 has normals, bit 25 has uvs
//   tri_emitter[i]  int     emitter index within the scene's emitter range, -1 if none
// BVH: 4-wide trees (one per scene, node indices world-wide), 128 B per node,
// children in SoA float4s:
//   node[8*n+0..5] = lo.x, hi.x, lo.y, hi.y, lo.z, hi.z of children 0..3
//   node[8*n+6]    = child index per slot (int bits): device node, or first leaf triangle
//   node[8*n+7]    = triangle count per slot (int bits): 0 = inner node, > 0 = leaf,
//                    -1 = absent slot (box [+inf, +inf]^3, skipped by the traversal);
//                    slot 0 keeps its count in bits 0..23 (two's complement for -1) and
//                    the node's slot masks in bits 24..31: leaf slots, then inner slots << 4
// Traversal nodes: the same tree re-encoded for the traversal kernels in
// kTravNodeWords = 20 words (80 B, five 16-byte loads) per node, same indices:
//   w[0..2]   node origin (float): the lowest lower child plane per axis
//   w[3]      bytes 0..2: exponent e per axis (int8); byte 3: slot kinds, bits 0..3
//             "leaf" slots, bits 4..7 "inner" slots: leaf only = leaf of 1..4
//             triangles, inner only = inner node, both = leaf of more triangles,
//             neither = absent slot
//   w[4..7]   x planes as fp16 offsets h: lower planes of children 0..3 (w4 = c0 |
//             c1 << 16, w5 = c2 | c3 << 16), then upper planes (w6, w7); the plane
//             is origin + h * 2^e, rounded outward (lower planes down, upper up)
//   w[8..11]  y planes, w[12..15] z planes, as x
//   w[16..19] child words: inner node index; leaf of n <= 4 triangles: first
//             triangle | (n - 1) << 30; larger leaf: index of its (first, count)
//             pair in the big-leaf table; absent: -1
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace sspt {

constexpr int kInfoMaterialMask = 0xFFFFFF;
constexpr int kInfoHasNormals = 1 << 24;
constexpr int kInfoHasUvs = 1 << 25;

// Depth bound of the binary SAH build.
constexpr int kMaxBvhDepth = 40;
// Nodes of one world: the traversal forms 32-bit byte offsets into the node array.
constexpr uint32_t kMaxNodes = 1u << 25;
// Traversal stack per lane: the top kStackLds entries live in LDS, the rest in
// a private (scratch) array; scene creation rejects trees that need more than
// kStackTotal entries.
#ifndef SSPT_STACK_LDS
#define SSPT_STACK_LDS 8
#endif
constexpr int kStackLds = SSPT_STACK_LDS;
constexpr int kStackTotal = 64;
constexpr int kTravNodeWords = 20;
#ifndef SSPT_TRACE_BLOCK
#define SSPT_TRACE_BLOCK 128
#endif
constexpr int kTraceBlock = SSPT_TRACE_BLOCK;  // threads per traversal block
// Leaves a lane may hold back for a later triangle phase (a power of two, at
// least 4: one node step adds up to four). They sit in LDS after the lane's
// stack entries, so a traversal kernel's node array holds kStackNodeLds ints.
#ifndef SSPT_PEND_CAP
#define SSPT_PEND_CAP 8
#endif
constexpr int kPendCap = SSPT_PEND_CAP;
static_assert(kPendCap >= 4 && (kPendCap & (kPendCap - 1)) == 0, "kPendCap: a power of two >= 4");
// A wave runs a triangle step once this percentage of its live lanes have a leaf queued.
#ifndef SSPT_PEND_PERCENT
#define SSPT_PEND_PERCENT 80
#endif
constexpr int kPendPercent = SSPT_PEND_PERCENT;
// The same for any-hit (shadow) rays: a queued leaf may end the ray.
#ifndef SSPT_PEND_PERCENT_ANY
#define SSPT_PEND_PERCENT_ANY 80
#endif
constexpr int kPendPercentAny = SSPT_PEND_PERCENT_ANY;
constexpr int kStackNodeLds = (kStackLds + kPendCap) * kTraceBlock;

// Queue entries pack (scene, slot): slot in the low 24 bits, scene above.
constexpr int kSlotBits = 24;
constexpr uint32_t kSlotMask = (1u << kSlotBits) - 1u;
constexpr int kMaxScenes = 128;
constexpr uint32_t kMaxSlots = 1u << kSlotBits;

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

constexpr int kEmitterGlossy = 1;  // DevEmitter::kind of a GlossyEmitter (0 = DiffuseEmitter)

struct DevEmitter {  // 32 B
    int tri;  // leaf-order triangle index (world-wide)
    float r, g, b;
    float inv_area;
    int kind;        // sspt_emitter_kind
    float exponent;  // glossy lobe exponent
    float norm;      // glossy: (exponent + 1) / (2 pi)
};

struct DevCamera {
    float v2c[16];  // view -> camera (Matrix4x4 row-major, row vectors)
    float c2w[16];  // camera -> world
    float w2c[16];  // world -> camera (PerspectiveCamera.SampleResponse)
    float c2v[16];  // camera -> view (the projection)
    float pos[3];
    float image_plane;  // distance to the image plane of unit-area pixels
    int width, height;
};

// Per-scene record. The render settings and the path range are refreshed by
// every render call.
struct SceneRec {
    DevCamera cam;
    int root;            // root node of the scene's BVH
    int emitter_off;     // first emitter of the scene in the emitter pool
    int num_emitters;
    int has_env;
    int env_w, env_h;
    uint32_t env_rgb_off;  // float4 index into the environment texel pool
    uint32_t env_row_off;  // float index into the CDF pool (env_h entries)
    uint32_t env_col_off;  // float index into the CDF pool (env_h * env_w entries)
    uint32_t env_guide_row_off;  // int index into the guide pool: env_guide_row_g + 1 entries
    uint32_t env_guide_col_off;  // env_h tables of env_guide_col_g + 1 entries
    int env_guide_row_g, env_guide_col_g;  // guide buckets (powers of two)
    float scene_center[3], scene_radius;   // Background.SceneCenter / SceneRadius (bidirectional light paths)
    float abs_max;         // largest vertex coordinate magnitude (ray slack scale)
    uint32_t num_pixels;
    // settings (PathTracerBase fields)
    uint32_t base_seed;
    int max_depth, min_depth, num_shadow_rays, enable_bsdf_di;
    int enable_denoiser;   // write the albedo / normal layers at primary hits
    // current render
    uint32_t first_iter;   // iteration index of the scene's first path in this render
    float adj_scale;       // adjoint renders: 1 / number of iterations (d image / d sample)
    uint64_t path_off;     // first world-wide path index of the scene in this render
    uint64_t fb_off;       // float offset of the scene's W*H*3 framebuffer
};

struct DevScene {
    const float4* nodes;      // the builders' 4-wide float tree (read back by sspt_scene_read_bvh)
    const uint32_t* qnodes;   // the same tree as traversal nodes
    const int* big_leaves;    // (first triangle, count) of the leaves of more than four triangles
    const float4* tri_pos;
    const float4* tri_aux;
    const int* tri_emitter;
    const DevMaterial* materials;
    const DevTexture* textures;
    const float* texels;
    const DevEmitter* emitters;
    const float4* env_rgb;  // pooled environment texels (w unused)
    const float* env_cdf;   // pooled row and column CDFs
    const int* env_guide;   // pooled CDF guide tables: guide[j] = lower bound of j / G in the CDF
    const SceneRec* recs;
    const uint64_t* path_offsets;  // num_scenes + 1 world-wide path offsets of this render
    const float* adj_image;        // adjoint renders: d loss / d image, framebuffer layout
    float* grad;                   // adjoint renders: d loss / d texel, texel-pool layout
    float* layers;                 // albedo layers (framebuffer layout), then normal layers at + layer_stride
    uint64_t layer_stride;
    // FrameBuffer outlier tracking: per pixel of the framebuffer pool, num_outliers
    // keys (average estimate, iteration) of the largest samples so far; null = off
    unsigned long long* outliers;
    int num_outliers;
    int num_scenes;
    int num_nodes;
    // set by the host before a render: every record agrees on what the shade
    // kernel reads (settings, emitters, environment), as in a world of one scene
    // or a view set rendered with one setting
    int recs_alike;
};

// Device-side per-wave work: the wave covers world-wide paths
// [start, start + count); in running-mean mode it covers whole iterations
// first_iter .. first_iter + iters - 1 of a single scene, or, with iters = 0,
// at most one iteration's worth of paths whose iterations count from first_iter.
struct WaveDesc {
    uint64_t start;
    uint32_t count;
    uint32_t first_iter;
    uint32_t iters;
    uint32_t pad;
};

// Device-side counters. Extension queue b and the shadow rays of the shade
// pass that fills it share the 64-bit word at counters + 2b (queue length in
// the low half, shadow rays in the high half), so one 64-bit atomic appends
// both.
enum Counter {
    kCntQueue0 = 0,   // extension queue 0 length
    kCntShadow0 = 1,  // shadow rays appended with queue 0
    kCntQueue1 = 2,   // extension queue 1 length
    kCntShadow1 = 3,  // shadow rays appended with queue 1
    kCntWave = 4,     // index of the wave being rendered
    kCntNum = 5
};
enum Stat {
    kStatRays = 0,
    kStatShadow = 1,        // next-event attempts (IsOccluded / LeavesScene calls)
    kStatOccluded = 2,      // traced shadow rays that were blocked
    kStatHits = 3,
    kStatNan = 4,
    kStatShadowTraced = 5,  // shadow rays traced (non-zero contribution)
    kStatNum = 6
};
// Ray statistics are striped over kStatStripes 128-byte lines: the waves of a
// launch add to different addresses (device-wide atomics on one address
// serialise across the XCDs), and a total is the sum over the stripes.
constexpr int kStatStripes = 256;
constexpr int kStatStride = 16;  // u64 per stripe (128 B)
constexpr unsigned long kStatBytes = 8ul * kStatStripes * kStatStride;


}  // namespace sspt
