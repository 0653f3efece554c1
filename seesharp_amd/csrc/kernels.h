// kernels.h -- launch interface of the wavefront path-tracer kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "scene_types.h"

namespace sspt {

// Wavefront buffers, all in HBM. P = path slots of one wave.
struct Wave {
    uint32_t capacity;     // P
    // extension queues (double buffered): origin+tmin, direction+tmax, (scene << 24 | slot),
    // and the live path state travelling with its ray (read and written in queue order)
    float4* q_o[2];
    float4* q_d[2];
    uint32_t* q_slot[2];
    uint32_t* q_rng[2];  // RNG state
    float4* q_thr[2];    // prefix weight rgb, previous pdf
    float4* q_prev[2];   // previous hit position, depth (int bits)
    float4* q_L[2];      // radiance from emitter / background hits so far
    float4* hits;          // (t, u, v, leaf triangle index bits), per queue entry
    // shadow queue: origin+tmin, direction+tmax, contribution rgb + (scene << 25 | slot << 1 | kind)
    float4* s_o;
    float4* s_d;
    float4* s_c;
    // per slot: written once, when the path ends (st_L, st_prev) or by other kernels
    uint32_t* st_pix;  // (scene << 24 | pixel)
    float4* st_prev;   // .w: the path's last depth (int bits)
    float4* st_L;      // radiance from emitter / background hits
    float4* st_Le;     // environment next-event contributions
    float4* st_La;     // area-light next-event contributions
    float4* st_alb;    // primary-hit albedo (scenes with denoiser layers)
    float4* st_nrm;    // primary-hit shading normal
    // adjoint renders: per vertex (index depth-1) and slot, [depth * capacity + slot]
    float4* ad_dlog;   // 3 float4 per record: d log f / d(base r,g,b, roughness) per colour channel
    float4* ad_A;      // radiance accumulated after the vertex's next-event rays resolved
    int2* ad_tex;      // float offsets of the vertex's albedo / roughness texels (-1 = none)
    uint32_t ad_depth; // records allocated per slot
    float4* s_g;       // per shadow ray: gradient of its contribution w.r.t. the vertex's 4 parameters
                       // (until k_shade_grad: the ray's local direction and queue entry)
    int2* s_t;         // per shadow ray: texel offsets of the vertex
    uint32_t* counters;            // kCntNum
    unsigned long long* stats;     // kStatStripes stripes of kStatStride (kStatNum used)
    const WaveDesc* waves;         // per-wave work of the current render
};

// Grid sizes of the queue kernels. Each kernel reads its queue's length on the
// device and walks the queue with a grid-stride loop. A grid size of 0 means
// one block per chunk of the longest the queue can be in this render
// (max_paths, max_shadow): blocks past the device-side length exit at once and
// the hardware dispatcher balances the rest block by block. A non-zero size
// gives persistent blocks that stride over the queue.
struct LaunchConfig {
    uint32_t trace_blocks;   // blocks of trace_block threads for extend / shadow (0 = full grid)
    uint32_t shade_blocks;   // blocks of shade_block threads for shade, 256 for generate / accumulate (0 = full grid)
    uint32_t trace_block = 0;  // threads per traversal block: 64, 128 or 256 (0 = 128)
    uint32_t shade_block = 0;  // threads per forward shade block: 128, 256 or 512 (0 = 256)
    uint32_t max_paths;      // largest wave of the render (paths)
    uint32_t max_shadow;     // largest shadow queue of the render (rays)
    int max_nsr;             // most shadow rays per kind and vertex over the render's scenes
};

void launch_wave_begin(hipStream_t s, const Wave& w);
void launch_generate(hipStream_t s, const DevScene& sc, const Wave& w, const LaunchConfig& lc);
// the single-kernel path tracer: a whole wave's paths, one thread per path
void launch_path(hipStream_t s, const DevScene& sc, const Wave& w, const LaunchConfig& lc);
void launch_bounce_reset(hipStream_t s, const Wave& w, int buf);
void launch_extend(hipStream_t s, const DevScene& sc, const Wave& w, const LaunchConfig& lc, int buf);
// launch_generate and launch_extend(buf 0) in one kernel
void launch_extend_primary(hipStream_t s, const DevScene& sc, const Wave& w, const LaunchConfig& lc);
void launch_shade(hipStream_t s, const DevScene& sc, const Wave& w, const LaunchConfig& lc, int buf);
// the shadow rays the shade pass appended with extension queue `slot`
void launch_shadow(hipStream_t s, const DevScene& sc, const Wave& w, const LaunchConfig& lc, int slot);
// running mean of one scene: waves of whole iterations, or (WaveDesc.iters = 0)
// waves of at most one iteration's pixel count
void launch_accumulate_mean(hipStream_t s, const Wave& w, float* image, uint32_t num_pixels, float* albedo,
                            float* normal, unsigned long long* outliers, int num_outliers);
// per-scene sums of a batch (any wave layout)
void launch_accumulate_sum(hipStream_t s, const DevScene& sc, const Wave& w, const LaunchConfig& lc,
                           float* image_pool);
void launch_wave_end(hipStream_t s, const Wave& w);
// n running-mean wave descriptors of a single-scene render written on the device
// (whole = pixel count for waves of whole iterations, 0 for partial-iteration waves)
void launch_fill_waves(hipStream_t s, WaveDesc* descs, uint32_t n, uint64_t first_path, uint32_t per_wave,
                       uint64_t total, uint32_t first_iter, uint32_t whole);
// adjoint (path replay) variants
void launch_shade_adjoint(hipStream_t s, const DevScene& sc, const Wave& w, const LaunchConfig& lc, int buf);
// The adjoint's dual-number material derivatives of the bounce just shaded (before its trace).
void launch_shade_grad(hipStream_t s, const DevScene& sc, const Wave& w, const LaunchConfig& lc, int buf);
void launch_shadow_adjoint(hipStream_t s, const DevScene& sc, const Wave& w, const LaunchConfig& lc, int slot);
// extension queue `buf` and the shadow queue in one grid (forward or adjoint shadow)
void launch_trace(hipStream_t s, const DevScene& sc, const Wave& w, const LaunchConfig& lc, int buf, bool adjoint);
void launch_adjoint_finalize(hipStream_t s, const DevScene& sc, const Wave& w, const LaunchConfig& lc);
void launch_scale(hipStream_t s, float* data, uint64_t n, float factor);
// Blocks per CU that fit at once (register / LDS limits) for the trace kernels
// (extend, shadow) and the shade kernel of the given block sizes; 0 if the query fails.
void resident_blocks_per_cu(uint32_t trace_block, uint32_t shade_block, int* trace, int* shade);
// Whether the block sizes are prebuilt ones (0 = the default size).
bool valid_block_shape(uint32_t trace_block, uint32_t shade_block);

// Ray queries on scene `scene` of the world: rays are 8 floats (ox,oy,oz,tmin,dx,dy,dz,tmax).
void launch_trace_closest(hipStream_t s, const DevScene& sc, int scene, const float* rays, uint64_t n,
                          float4* hits_out);
void launch_trace_occluded(hipStream_t s, const DevScene& sc, int scene, const float* rays, uint64_t n,
                           uint8_t* occ);
void launch_primary_rays(hipStream_t s, const DevScene& sc, int scene, uint32_t num_pixels, uint32_t base_seed,
                         uint32_t iteration, float* rays);

// DebugVisualizer: eye-light shading of the primary hits of iterations
// first..first+num-1 of scene record `scene` into its running-mean image.
// One iteration of the light tracer: num_paths light paths seeded
// (base_seed, path, iteration), splats / n added to `image` (the scene's framebuffer).
struct LightParams {
    uint32_t base_seed, iteration, num_paths;
    int max_depth, min_depth;
    float num_paths_f;  // NumLightPaths as a float
    float n;            // FrameBuffer.CurIteration
};
void launch_light_paths(hipStream_t s, const DevScene& sc, int scene, const LightParams& p, float* image,
                        unsigned long long* stats);
// One iteration of ClassicBidir over the pixels of scene record `scene`
// (num_paths = width * height light paths and camera paths).
struct BidirParams {
    uint32_t seed_camera, seed_light, iteration, num_paths;
    int max_depth, min_depth, num_shadow_rays;
    int connections, light_tracer;  // EnableConnections, EnableLightTracer
    float num_paths_f;              // NumLightPaths as a float
    float n;                        // FrameBuffer.CurIteration
    // The light-path cache, one level per vertex depth: vertex k of path p at
    // [k * num_paths + p] (k < max_depth).
    float4* lv_a;   // prefix weight rgb, PdfFromAncestor
    float4* lv_b;   // barycentric u, v, PdfReverseAncestor, triangle (int bits)
    int* lv_len;    // vertices of path p (0: none)
    float* lv_ne;   // next-event pdf of path p's emitter point (PdfNextEventAncestor)
    // Camera vertex k (depth k + 1) of pixel p at [k * num_pixels + p]:
    // PdfFromAncestor, PdfToAncestor.
    float2* cam;
    uint32_t num_pixels;  // camera paths (ClassicBidir: num_paths)
    int env;              // the scene has a background (the kernels' instantiation with its code)
    // ClassicBidir's camera paths as a wavefront (per pixel p; records at [p * bc_rec + r]):
    // live paths' state, hit, the vertex's emitter term and prefix, shadow records
    uint32_t* bc_rng;
    float4* bc_a;         // prefix rgb, pdf of the last direction
    float4* bc_b;         // approximate throughput rgb, depth (int bits)
    float4* bc_o;         // ray origin, tmin
    float4* bc_d;         // ray direction
    float4* bc_pp;        // previous vertex position
    float4* bc_pn;        // previous vertex normal
    float4* bc_L;         // the path's sum
    float4* bc_hit;       // u, v, t, triangle (int bits; -1 = miss)
    float4* bc_v0;        // the vertex's prefix-weighted emitter / background term, connections (bits)
    float4* bc_v1;        // the vertex's prefix, next-event records (bits)
    float4* bc_r0;        // record: segment start (or ray origin), tmin of a ray to infinity
    float4* bc_r1;        // record: segment end (or direction), w = 1 for a ray to infinity
    float4* bc_rc;        // record: unweighted term (mis * w), visibility
    uint32_t* bc_list0;   // live paths of even bounces
    uint32_t* bc_list1;   // of odd bounces
    uint32_t* bc_shadow;  // record slots to test
    uint32_t* bc_count;   // [0], [1]: list sizes, [2]: records to test
    int bc_rec;           // records per path and bounce: max_depth - 1 connections + num_shadow_rays
    // VertexCacheBidir / VertexConnectionAndMerging
    int mode;             // kModeClassic, kModeVertexCache, kModeVcm
    int num_connections, hitting, merging, merge_primary, max_photons, correl;
    int* cum;             // inclusive prefix sums of lv_len (PathCache.cumPathLen)
    // photon map: a hash grid of cells of size `cell`; bucket b holds
    // photons[bucket_start[b] .. bucket_start[b + 1]) (position, global vertex index bits)
    float cell;
    uint32_t table_mask;
    uint32_t* bucket_start;
    uint32_t* bucket_fill;
    float4* photons;
};
constexpr int kModeClassic = 0, kModeVertexCache = 1, kModeVcm = 2;
constexpr int kMaxPhotons = 32;  // largest MaxNumPhotons of a merge
// TraceLightPaths (+ SplatLightVertices if light_tracer), then the camera paths.
void launch_bidir_light(hipStream_t s, const DevScene& sc, int scene, const BidirParams& p, float* image,
                        unsigned long long* stats);
// The camera paths as a wavefront: per bounce a trace, a shade, a shadow and a
// gather kernel.
void launch_bidir_camera_wave(hipStream_t s, const DevScene& sc, int scene, const BidirParams& p, float* image,
                              unsigned long long* stats);
// VertexCacheBidir / VCM: the light paths' global vertex indices (p.cum), the
// photon grid (VCM with merging) and the camera paths. `scan_tmp` holds
// scan_tmp_words(max(num_paths, table size)) ints.
uint32_t scan_tmp_words(uint32_t n);
void launch_vcm_prepare(hipStream_t s, const DevScene& sc, int scene, const BidirParams& p, int* scan_tmp);
void launch_vcm_camera(hipStream_t s, const DevScene& sc, int scene, const BidirParams& p, float* image,
                       unsigned long long* stats);
void launch_debug_visualize(hipStream_t s, const DevScene& sc, int scene, uint32_t num_pixels, uint32_t base_seed,
                            uint32_t first, uint32_t num, float* image);

// One node of a replayed path (the layout of sspt_path_node).
struct PathNode {
    int kind, ancestor;
    float position[3], emission[3], mis_weight, scatter_weight[3], survival_probability, pdf, bsdf_cos[3],
        prefix_weight[3];
    uint32_t flags;
};
// ReplayPixel of pixel `pix`, `iteration` of scene record `scene` (settings
// already in the record): writes up to `cap` nodes, the node count and the
// path's RGB estimate to device memory.
void launch_replay(hipStream_t s, const DevScene& sc, int scene, uint32_t pix, uint32_t iteration, PathNode* nodes,
                   uint32_t cap, uint32_t* count, float* estimate);

}  // namespace sspt
