/*
 * sspt.h -- C ABI of the MI355X-native SeeSharp path tracer (libsspt.so).
 *
 * The library is the GPU replacement for the CPU hot path of SeeSharp's
 * unidirectional `PathTracer` and of the TinyEmbree ray tracer underneath it.
 * Every entry point below names the reference interface it replaces.
 *
 * Conventions
 *   - Every call returns an sspt_status (0 = OK); nothing throws across the ABI.
 *     The message of the last failure is per context: sspt_last_error().
 *   - Host arrays passed in are owned by the caller and are copied; device
 *     memory is owned by the context / scene handles.
 *   - One HIP stream per context. A context is not thread-safe, mirroring
 *     `Integrator.Render`, which renders one scene at a time
 *     (SeeSharp/Integrators/PathTracer.cs:196-200).
 *   - Matrices use System.Numerics.Matrix4x4 layout: row-major M11..M44 with
 *     row vectors (v' = v * M), as SeeSharp/Cameras/Camera.cs:34-42 stores them.
 */
#ifndef SSPT_H
#define SSPT_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SSPT_ABI_VERSION 8

typedef enum sspt_status {
    SSPT_OK = 0,
    SSPT_ERR_INVALID_ARGUMENT = 1,
    /* Scene.Prepare on an invalid scene (SeeSharp/Scene.cs:104-105,170-197) */
    SSPT_ERR_INVALID_SCENE = 2,
    SSPT_ERR_DEVICE = 3,
    SSPT_ERR_OUT_OF_MEMORY = 4,
    SSPT_ERR_UNSUPPORTED = 5,
    /* Triangle with zero area used as an emitter (SeeSharp/Geometry/Triangle.cs:32-33) */
    SSPT_ERR_ZERO_AREA_EMITTER = 6
} sspt_status;

typedef struct sspt_ctx sspt_ctx;
typedef struct sspt_scene sspt_scene;

/* One `Mesh` (SeeSharp/Geometry/Mesh.cs:6-30). PrimId stays mesh-local. */
typedef struct sspt_mesh_desc {
    uint32_t num_vertices;
    const float* positions;   /* 3 * num_vertices */
    const float* normals;     /* 3 * num_vertices, or NULL: shading normal = face normal */
    const float* uvs;         /* 2 * num_vertices, or NULL: texture coordinates (0,0) */
    uint32_t num_triangles;
    const int32_t* indices;   /* 3 * num_triangles */
    int32_t material;         /* index into sspt_scene_desc.materials, -1 = none (invalid) */
} sspt_mesh_desc;

typedef enum sspt_material_kind {
    SSPT_MATERIAL_DIFFUSE = 0, /* DiffuseMaterial (SeeSharp/Shading/Materials/DiffuseMaterial.cs:6) */
    SSPT_MATERIAL_GENERIC = 1  /* GenericMaterial (SeeSharp/Shading/Materials/GenericMaterial.cs:3) */
} sspt_material_kind;

/* Tagged union of DiffuseMaterial.Parameters (DiffuseMaterial.cs:193) and
 * GenericMaterial.Parameters (GenericMaterial.cs:22-60). Texture ids index
 * sspt_scene_desc.textures; -1 = the constant value. */
typedef struct sspt_material_desc {
    int32_t kind;
    float base_color[3];
    int32_t base_color_texture;     /* RGB texture or -1 */
    float roughness;
    int32_t roughness_texture;      /* mono texture or -1 */
    float metallic;
    float specular_tint_strength;
    float anisotropic;
    float specular_transmittance;
    float index_of_refraction;
    int32_t transmitter;            /* DiffuseMaterial.Parameters.Transmitter */
} sspt_material_desc;

/* TextureRgb / TextureMono image (SeeSharp/Images/ImageTexture.cs:6-50). */
typedef struct sspt_texture_desc {
    uint32_t width, height;
    uint32_t channels;     /* 3 (RGB) or 1 (mono) */
    int32_t border;        /* 0 = Repeat (default), 1 = Clamp */
    const float* texels;   /* row-major, channels floats per texel */
} sspt_texture_desc;

typedef enum sspt_emitter_kind {
    /* DiffuseEmitter (SeeSharp/Shading/Emitters/DiffuseEmitter.cs:6-37): radiance
     * on the side the shading normal faces */
    SSPT_EMITTER_DIFFUSE = 0,
    /* GlossyEmitter (SeeSharp/Shading/Emitters/GlossyEmitter.cs:7-40): radiance *
     * cos^exponent * (exponent + 1) / (2 pi) on that side */
    SSPT_EMITTER_GLOSSY = 1
} sspt_emitter_kind;

/* One emissive triangle, listed in Scene.Emitters order: next-event estimation
 * indexes that order. */
typedef struct sspt_emitter_desc {
    uint32_t mesh;
    uint32_t face;
    float radiance[3];
    int32_t kind;          /* sspt_emitter_kind */
    float exponent;        /* GlossyEmitter only: finite and > -1 */
} sspt_emitter_desc;

/* EnvironmentMap lat-long image (SeeSharp/Shading/Background/EnvironmentMap.cs:13-189). */
typedef struct sspt_envmap_desc {
    uint32_t width, height;
    const float* rgb;      /* row-major RGB */
} sspt_envmap_desc;

/* PerspectiveCamera without depth of field (SeeSharp/Cameras/PerspectiveCamera.cs:6-52). */
typedef struct sspt_camera_desc {
    float world_to_camera[16];
    float vertical_fov_degrees;
} sspt_camera_desc;

/* The `Scene` (SeeSharp/Scene.cs:6-60), flattened in Meshes / Emitters order. */
typedef struct sspt_scene_desc {
    uint32_t num_meshes;
    const sspt_mesh_desc* meshes;
    uint32_t num_materials;
    const sspt_material_desc* materials;
    uint32_t num_textures;
    const sspt_texture_desc* textures;
    uint32_t num_emitters;
    const sspt_emitter_desc* emitters;
    const sspt_envmap_desc* background;  /* NULL = no background */
    sspt_camera_desc camera;
    uint32_t width, height;              /* FrameBuffer resolution */
} sspt_scene_desc;

/* Public fields of PathTracerBase (SeeSharp/Integrators/PathTracer.cs:16-49) and
 * Integrator (SeeSharp/Integrators/Integrator.cs:10-16). */
typedef struct sspt_pt_settings {
    uint32_t base_seed;          /* BaseSeed, default 0xC030114 */
    int32_t total_spp;           /* TotalSpp */
    int32_t max_depth;           /* MaxDepth */
    int32_t min_depth;           /* MinDepth */
    int32_t num_shadow_rays;     /* NumShadowRays */
    int32_t enable_bsdf_di;      /* EnableBsdfDI */
    int64_t max_render_time_ms;  /* MaximumRenderTimeMs, < 0 = unlimited. With a budget sspt_render
                                    renders one iteration per wave and, before each, stops once the
                                    whole milliseconds spent so far plus their average per iteration
                                    exceed it (RenderTimer, PathTracer.cs:221-227); stats.num_iterations
                                    says how many it rendered */
    int32_t enable_denoiser;     /* EnableDenoiser: writes the "albedo" and "normal" layers at primary
                                    hits (DenoiseBuffers.LogPrimaryHit, PathTracer.cs:336-339); the
                                    OIDN denoiser itself is not run */
} sspt_pt_settings;

/* TinyEmbree RayTracerStats as stored by PathTracer.cs:257: num_rays counts
 * closest-hit traces, num_shadow_rays every IsOccluded / LeavesScene call,
 * num_ray_hits the traces that hit. Shadow rays whose contribution is exactly
 * zero are not traced: num_shadow_rays_traced counts the traced ones and
 * num_occluded the traced ones that were blocked. */
typedef struct sspt_ray_stats {
    uint64_t num_rays;
    uint64_t num_shadow_rays;
    uint64_t num_occluded;
    uint64_t num_ray_hits;
    uint64_t num_shadow_rays_traced;
    /* Iterations the call rendered: sspt_render renders all it was asked for
     * unless max_render_time_ms stopped it earlier (the FrameBuffer's
     * CurIteration / NumIterations); sspt_render_adjoint the iterations it
     * replayed; the batch calls the largest total_spp of their scenes. */
    uint64_t num_iterations;
} sspt_ray_stats;

/* Closest-hit record of TinyEmbree `Hit` for the trace entry points. */
typedef struct sspt_hit {
    float t;             /* Hit.Distance, +inf on a miss */
    float u, v;          /* Hit.BarycentricCoords (weights of vertex 1 and 2) */
    int32_t prim;        /* global triangle id (mesh offset + PrimId), -1 on a miss */
} sspt_hit;

/* ---- context ---------------------------------------------------------- */

int32_t sspt_abi_version(void);

/* Creates a context on HIP device `device` with its own stream. */
sspt_status sspt_ctx_create(int32_t device, sspt_ctx** out_ctx);
void sspt_ctx_destroy(sspt_ctx* ctx);
/* Message of the last failed call on this context ("" if none). */
const char* sspt_last_error(const sspt_ctx* ctx);

/* ---- scene: replaces Scene.Prepare (SeeSharp/Scene.cs:103-164), i.e. the
 *      TinyEmbree Raytracer.AddMesh/CommitScene BVH build (:108-113), the
 *      emitter maps (:147-163) and the environment-map sampling grid
 *      (SeeSharp/Shading/Background/EnvironmentMap.cs:139-156). --------- */
sspt_status sspt_scene_create(sspt_ctx* ctx, const sspt_scene_desc* desc,
                              sspt_scene** out_scene);
void sspt_scene_destroy(sspt_scene* scene);
/* Triangle count, emitter count, BVH node count and device bytes held. */
sspt_status sspt_scene_info(const sspt_scene* scene, uint64_t* num_triangles,
                            uint64_t* num_emitters, uint64_t* num_bvh_nodes,
                            uint64_t* device_bytes);

/* BVH builder of the scenes a context creates (the CommitScene step,
 * SeeSharp/Scene.cs:108-113): binned SAH on the device (default) or on the
 * host. Both make the same tree: the same nodes, boxes and leaf order (a box
 * bound may be -0 where the other has +0). */
typedef enum sspt_bvh_builder {
    SSPT_BVH_DEVICE = 0,
    SSPT_BVH_HOST = 1
} sspt_bvh_builder;
sspt_status sspt_set_bvh_builder(sspt_ctx* ctx, int32_t builder);
sspt_status sspt_get_bvh_builder(const sspt_ctx* ctx, int32_t* builder);
/* Triangles per BVH leaf at most (1..16, default 2) for the scenes a context
 * creates, both builders. No counterpart in the reference (Embree picks its own
 * leaves); it changes the tree, never a hit: the closest hit is decided by
 * distance, then by the lower triangle id. */
sspt_status sspt_set_bvh_leaf_size(sspt_ctx* ctx, int32_t max_triangles);
sspt_status sspt_get_bvh_leaf_size(const sspt_ctx* ctx, int32_t* max_triangles);
/* The builder a scene was built with, the summed duration of its build
 * kernels (0 for the host builder) and the wall time of its BVH builds. */
sspt_status sspt_scene_bvh_build_time(const sspt_scene* scene, int32_t* builder,
                                      double* kernel_ms, double* wall_ms);
/* Copies the scene's 4-wide BVH as the builder made it (the traversal reads
 * the same tree, node for node, with its child planes re-encoded as fp16
 * offsets rounded outward): max_nodes >= num_bvh_nodes
 * nodes of 32 floats (per axis the 4 child box minima then maxima, then 4
 * child indices and 4 counts as int32, the top byte of the first count
 * holding the node's leaf and inner slot masks), and max_tris >= num_triangles
 * global triangle ids, one per leaf position. */
sspt_status sspt_scene_read_bvh(sspt_ctx* ctx, sspt_scene* scene, float* nodes, uint64_t max_nodes,
                                int32_t* leaf_triangles, uint64_t max_tris);

/* ---- batching: many scenes (cameras, resolutions, triangle counts) in one
 *      device world, rendered in shared waves. No counterpart in the reference,
 *      which renders one Scene per Integrator.Render call; each scene of a batch
 *      keeps the semantics of its own PathTracer render. --------------------- */
/* Builds n (<= 128) scenes into one world (Scene.Prepare for each). */
sspt_status sspt_batch_create(sspt_ctx* ctx, const sspt_scene_desc* descs, uint32_t n,
                              sspt_scene** out_scene);
sspt_status sspt_scene_count(const sspt_scene* scene, uint32_t* n);
/* Renders settings[i].total_spp iterations (0 .. total_spp-1) of every scene i
 * from a cleared framebuffer; each framebuffer ends as the mean of its
 * samples (fp32 sums in arbitrary order, divided by total_spp). */
sspt_status sspt_render_batch(sspt_ctx* ctx, sspt_scene* scene, const sspt_pt_settings* settings,
                              sspt_ray_stats* stats);
/* A view set: num_views cameras on one scene. The world holds the scene's
 * geometry, BVH, materials, textures and emitters once and one record and
 * framebuffer per view (desc's own camera is unused; every view has desc's
 * width and height). Render it with sspt_render_batch (one settings entry per
 * view) and read view i with sspt_framebuffer_read_scene; adjoint renders of
 * all views add into the one set of texture gradients. No counterpart in the
 * reference (SeeSharp renders one camera per Scene). */
sspt_status sspt_views_create(sspt_ctx* ctx, const sspt_scene_desc* desc, const sspt_camera_desc* cameras,
                              uint32_t num_views, sspt_scene** out_scene);
/* Copies scene `index`'s W*H*3 framebuffer to host memory. */
sspt_status sspt_framebuffer_read_scene(sspt_ctx* ctx, sspt_scene* scene, uint32_t index,
                                        float* host_rgb);

/* ---- execution knobs (the autotuner's search space) ---------------------
 * Wave capacity in path slots (0 = default 16 Mi, the most a wave holds); grid sizes of the
 * traversal (extend, shadow) and shading kernels in blocks: 0 (the default)
 * launches one block per chunk of the longest its queue can be and lets the
 * hardware dispatcher balance the blocks, a non-zero size launches that many
 * persistent blocks striding over the queue; whether a wave is captured once
 * as a hipGraph and replayed (default on; off while kernel timing is
 * enabled). */
sspt_status sspt_set_wave_capacity(sspt_ctx* ctx, uint32_t slots);
sspt_status sspt_set_launch_config(sspt_ctx* ctx, uint32_t trace_blocks, uint32_t shade_blocks,
                                   int32_t use_graphs);
/* Execution mode of sspt_render / sspt_render_batch: the wavefront tracer
 * (default; generate / extend / shade / shadow kernels over queues), or the
 * plain single-kernel tracer (one thread follows one path to its end, shadow
 * rays traced where sampled): the same estimator and RNG draws, kept as the
 * per-path baseline. Adjoint renders need the wavefront. */
typedef enum sspt_exec_mode {
    SSPT_EXEC_WAVEFRONT = 0,
    SSPT_EXEC_SINGLE_KERNEL = 1
} sspt_exec_mode;
sspt_status sspt_set_execution_mode(sspt_ctx* ctx, int32_t mode);
/* Block shape: threads per block of the traversal kernels (64, 128 or 256)
 * and of the forward shade kernel (128, 256 or 512), 0 = the defaults (128,
 * 256); each is a prebuilt kernel variant. Grid sizes set with
 * sspt_set_launch_config count blocks of these sizes. Images do not depend
 * on the shape. */
sspt_status sspt_set_block_shape(sspt_ctx* ctx, uint32_t trace_block, uint32_t shade_block);
sspt_status sspt_get_block_shape(sspt_ctx* ctx, uint32_t* trace_block, uint32_t* shade_block);
/* Configured grid sizes (0 = full grid) and the blocks per CU of the
 * traversal and shading kernels (of the current block shape) that are
 * resident at once. */
sspt_status sspt_get_launch_config(sspt_ctx* ctx, uint32_t* trace_blocks, uint32_t* shade_blocks,
                                   uint32_t* trace_per_cu, uint32_t* shade_per_cu);

/* ---- framebuffer: the scene's FrameBuffer.Image running mean
 *      (SeeSharp/Images/FrameBuffer.cs:89-160). ------------------------- */
/* Zeroes the device image (FrameBuffer with CurIteration = 0). */
sspt_status sspt_framebuffer_clear(sspt_ctx* ctx, sspt_scene* scene);
/* Copies the W*H*3 fp32 running mean to host memory. */
sspt_status sspt_framebuffer_read(sspt_ctx* ctx, sspt_scene* scene, float* host_rgb);
/* The denoiser feature layers (DenoiseBuffers, SeeSharp/Integrators/Util/DenoiseBuffers.cs:10-32):
 * "albedo" = Material.GetScatterStrength and "normal" = the shading normal at the
 * primary hit, W*H*3 running means like the image (RgbLayer.Splat, Layer.OnStartIteration),
 * written by renders with enable_denoiser set. `index` selects the scene / view of
 * a batch (0 for a single scene). */
typedef enum sspt_layer { SSPT_LAYER_ALBEDO = 0, SSPT_LAYER_NORMAL = 1 } sspt_layer;
sspt_status sspt_framebuffer_read_layer(sspt_ctx* ctx, sspt_scene* scene, uint32_t index, int32_t layer,
                                        float* host_rgb);
/* FrameBuffer.NumOutliersToTrack and its OutlierReplayCache
 * (SeeSharp/Images/FrameBuffer.cs:29, 186-189;
 * SeeSharp/Integrators/Util/OutlierReplayCache.cs:9-22): per pixel, the n
 * samples with the largest average estimate ((R+G+B)/3 of the splatted value)
 * and the iterations that drew them; among equal weights the earliest
 * iterations stay. n <= SSPT_MAX_OUTLIERS applies to every scene / view of the
 * world; 0 (the default) turns tracking off. Tracking starts empty,
 * sspt_framebuffer_clear and sspt_render_batch empty it, sspt_render adds to
 * it. Non-finite samples are not tracked. */
#define SSPT_MAX_OUTLIERS 4
sspt_status sspt_framebuffer_track_outliers(sspt_ctx* ctx, sspt_scene* scene, uint32_t n);
/* Scene / view `index`'s tracked samples: W*H*n weights and iterations (either
 * may be NULL), per pixel in descending weight, the order FrameBuffer.WriteToFile
 * writes its "outlier-i" layers; weight 0 and iteration -1 where a pixel has
 * fewer than n samples. */
sspt_status sspt_framebuffer_read_outliers(sspt_ctx* ctx, sspt_scene* scene, uint32_t index, float* weights,
                                           int32_t* iterations);
/* Number of non-finite path estimates skipped so far (FrameBuffer.NaNWarnings count). */
sspt_status sspt_framebuffer_nan_count(sspt_ctx* ctx, sspt_scene* scene, uint64_t* count);

/* ---- rendering: replaces PathTracerBase.Render's iteration loop
 *      (SeeSharp/Integrators/PathTracer.cs:220-252), RenderPixel (:287-313)
 *      and EstimateIncidentRadiance (:315-382). -------------------------
 * Renders iterations first_iteration .. first_iteration+num_iterations-1
 * (one sample per pixel each, RNG seeded by (base_seed, row*W+col, iteration))
 * into the scene framebuffer with FrameBuffer running-mean semantics: at
 * iteration n (1-based) the image is scaled by (n-1)/n and the path estimate
 * divided by n is added. `stats` (may be NULL) receives the ray counters of
 * this call. Blocks until the work is done. */
sspt_status sspt_render(sspt_ctx* ctx, sspt_scene* scene, const sspt_pt_settings* settings,
                        uint32_t first_iteration, uint32_t num_iterations,
                        sspt_ray_stats* stats);

/* ---- allocation-free rendering. No counterpart in the reference: it is for
 *      callers that cannot allocate device memory (code under HIP graph capture,
 *      tight serving loops, harnesses that forbid hipMalloc). -----------------
 * sspt_render's semantics and results (the same images bit for bit, the same
 * ray counters) with every wave buffer in one file-scope device array of
 * SSPT_NOALLOC_SCRATCH_BYTES per GPU instead of device memory the context
 * allocates; the call makes no hipMalloc / hipFree, no graph and no event.
 * Waves hold as many path slots as fit the array next to a shadow queue of
 * 2 * num_shadow_rays entries per slot (about 172 000 slots at one shadow ray);
 * an image with more pixels is rendered in waves of part of an iteration.
 * Limits: num_shadow_rays small enough to leave 256 slots (up to about 2 700),
 * no time budget (max_render_time_ms must be < 0: SSPT_ERR_UNSUPPORTED).
 * The scene's framebuffer (allocated by sspt_scene_create) receives the image.
 * Every context on a GPU shares that GPU's array and each call leaves its last
 * contents behind: calls that overlap on one GPU (two threads, or two contexts
 * on one device) race and must be serialised by the caller. */
#define SSPT_NOALLOC_SCRATCH_BYTES (64u << 20)
sspt_status sspt_render_noalloc(sspt_ctx* ctx, sspt_scene* scene, const sspt_pt_settings* settings,
                                uint32_t first_iteration, uint32_t num_iterations,
                                sspt_ray_stats* stats);

/* ---- differentiation: path-replay adjoint of sspt_render. No counterpart in
 *      the reference. Given dLoss/dImage for the image of iterations
 *      first_iteration .. first_iteration+num_iterations-1, replays those paths
 *      from their seeds and adds dLoss/dtexel of every textured base colour and
 *      roughness to the scene's gradient buffers (BSDF values differentiated;
 *      sampling pdfs, MIS weights and Russian roulette held constant).
 *      dLoss/dImage may be host memory or device memory of the context's
 *      device (unified addressing: a loss computed on the GPU is passed by
 *      its device pointer and copied device to device). --- */
sspt_status sspt_render_adjoint(sspt_ctx* ctx, sspt_scene* scene, const sspt_pt_settings* settings,
                                const float* dloss_dimage, uint32_t first_iteration,
                                uint32_t num_iterations, sspt_ray_stats* stats);
/* The adjoint of sspt_render_batch for every scene / view of a world at once:
 * dloss_dimages (host or device memory) holds each scene's W*H*3 dLoss/dImage one after the other
 * (scene order); scene i replays iterations first_iteration ..
 * first_iteration + total_spp_i - 1. Gradients add into the world's texture
 * gradients (shared by all views of a view set). No counterpart in the
 * reference. */
sspt_status sspt_render_adjoint_batch(sspt_ctx* ctx, sspt_scene* scene, const sspt_pt_settings* settings,
                                      const float* dloss_dimages, uint32_t first_iteration,
                                      sspt_ray_stats* stats);
/* Zeroes every texture gradient of the scene. */
sspt_status sspt_gradient_clear(sspt_ctx* ctx, sspt_scene* scene);
/* Copies the gradient of texture `texture_index` (width*height*channels floats). */
sspt_status sspt_gradient_read(sspt_ctx* ctx, sspt_scene* scene, uint32_t texture_index, float* host);
/* Replaces the texels of texture `texture_index` (the differentiable parameters). */
sspt_status sspt_texture_write(sspt_ctx* ctx, sspt_scene* scene, uint32_t texture_index,
                               const float* host_texels);

/* ---- ray queries: TinyEmbree Raytracer.Trace (PathTracer.cs:319) and
 *      IsOccluded (:467) over the scene BVH, for callers and parity tests.
 * rays: n records of 8 floats (ox,oy,oz,tmin, dx,dy,dz,tmax). --------- */
sspt_status sspt_trace_closest(sspt_ctx* ctx, sspt_scene* scene, const float* host_rays,
                               uint64_t n, sspt_hit* host_hits);
sspt_status sspt_trace_occluded(sspt_ctx* ctx, sspt_scene* scene, const float* host_rays,
                                uint64_t n, uint8_t* host_occluded);

/* Primary rays of one iteration: RNG jitter + PerspectiveCamera.GenerateRay
 * (PathTracer.cs:289-291, PerspectiveCamera.cs:64-102) followed by the closest
 * hit. rays_out: W*H*8 floats as above (may be NULL); hits_out: W*H (may be NULL). */
sspt_status sspt_primary_hits(sspt_ctx* ctx, sspt_scene* scene, uint32_t base_seed,
                              uint32_t iteration, float* host_rays_out,
                              sspt_hit* host_hits_out);

/* ---- path replay: Integrator.ReplayPixel (SeeSharp/Integrators/Integrator.cs:28-29),
 *      PathTracerBase.ReplayPixel (SeeSharp/Integrators/PathTracer.cs:61-69). ----
 * Re-traces the one path of pixel (col, row) in `iteration` and returns its
 * PathGraph (SeeSharp/Integrators/Util/PathGraph.cs) as a node array in the
 * order RenderPixel / EstimateIncidentRadiance add the nodes (PathTracer.cs:303,
 * 326, 346-348, 447, 496), each node pointing at its ancestor. */
typedef enum sspt_node_kind {
    SSPT_NODE_CAMERA = 0,       /* the root: the primary ray's origin */
    SSPT_NODE_BSDF_SAMPLE = 1,  /* BSDFSampleNode: a surface hit of the path */
    SSPT_NODE_NEXT_EVENT = 2,   /* NextEventNode: an unoccluded next-event sample with a non-black contribution */
    SSPT_NODE_BACKGROUND = 3    /* BackgroundNode: the path left the scene at depth >= MinDepth */
} sspt_node_kind;
#define SSPT_NODE_HIT_EMITTER 1u    /* flags: a BSDF sample that hit an emitter (emission, mis_weight set) */
#define SSPT_NODE_TO_BACKGROUND 2u  /* flags: a next-event sample of the background */
typedef struct sspt_path_node {
    int32_t kind;                /* sspt_node_kind */
    int32_t ancestor;            /* index of the ancestor node, -1 for the root */
    float position[3];           /* directions (background, background next events): ancestor position + direction */
    float emission[3];           /* hit emitter / background radiance; next event: Emission */
    float mis_weight;
    float scatter_weight[3];     /* BSDF sample: PreviousScatterWeight (sample weight / survival probability) */
    float survival_probability;  /* BSDF sample: PreviousSurvivalProbability */
    float pdf;                   /* next event: Pdf */
    float bsdf_cos[3];           /* next event: BsdfTimesCosine */
    float prefix_weight[3];      /* next event: PrefixWeight */
    uint32_t flags;
} sspt_path_node;
/* Scene / view `index` of the world (0 for a single scene). Writes at most
 * `capacity` nodes; *num_nodes receives the path's node count (which may
 * exceed capacity: 2 + max_depth * (1 + 2 * num_shadow_rays) always suffices)
 * and estimate_rgb (may be NULL) the path's radiance estimate, the value the
 * render splats for this sample. */
sspt_status sspt_replay_pixel(sspt_ctx* ctx, sspt_scene* scene, uint32_t index, const sspt_pt_settings* settings,
                              uint32_t col, uint32_t row, uint32_t iteration, sspt_path_node* host_nodes,
                              uint32_t capacity, uint32_t* num_nodes, float* estimate_rgb);

/* ---- DebugVisualizer (SeeSharp/Integrators/DebugVisualizer.cs:6-59) --------
 * Renders iterations first_iteration .. first_iteration+num_iterations-1 of the
 * grayscale "eye light" view of scene / view `index`: per sample the jittered
 * primary ray (RNG seeded by (base_seed, row*W+col, iteration)), |cos| between
 * the hit's face normal and the direction back to the camera, 0 on a miss,
 * into the framebuffer with running-mean semantics (clear it first for a new
 * image). */
sspt_status sspt_render_debug(sspt_ctx* ctx, sspt_scene* scene, uint32_t index, uint32_t base_seed,
                              uint32_t first_iteration, uint32_t num_iterations);

/* ---- light tracing: BidirBase.Render (SeeSharp/Integrators/Bidir/BidirBase.cs:84-146)
 *      of a VertexCacheBidir (SeeSharp/Integrators/Bidir/VertexCacheBidir.cs) set
 *      up as a pure light tracer: NumConnections = 0, NumShadowRays = 0,
 *      EnableHitting = false, EnableLightTracer = true. ---------------------
 * Each iteration traces num_light_paths paths from the emitters, seeded
 * RNG(base_seed_light, path, iteration) (TraceLightPaths / TraceEmitterPath,
 * BidirBase.Light.cs:192-277: SelectLight, the emitter's SampleRay
 * (DiffuseEmitter.cs:44-64, GlossyEmitter.cs:47-70), RandomWalk.StartFromEmitter
 * and ContinueWalk, RandomWalk.cs:82-90,150-217), and connects every vertex
 * of every path to the camera (ConnectLightVertexToCamera,
 * BidirBase.Light.cs:35-92, PerspectiveCamera.SampleResponse). With the other
 * techniques off the light tracer's MIS weight is 1 and the camera paths add
 * nothing, so they are not traced. The splats go into the framebuffer of the
 * (single) scene with running-mean semantics: each iteration n first scales
 * the image by (n - 1) / n, then adds every finite splat / n in an
 * unspecified order (fp32 atomics); non-finite splats are counted
 * (sspt_framebuffer_nan_count). stats count the light paths' closest-hit
 * traces and their camera connections' occlusion tests. With a background
 * (EnvironmentMap) this configuration also traces camera paths, which add the
 * background's radiance under EmitterHitMis, and light paths start from it as
 * in sspt_render_bidir: the call is then sspt_render_vcm with
 * SSPT_VERTEX_CACHE_BIDIR, no connections, next events or emitter hits, and
 * BidirBase's default base_seed_camera 0xC030114. Scenes without emitters or
 * background, and batches, are SSPT_ERR_UNSUPPORTED. */
typedef struct sspt_lt_settings {
    uint32_t base_seed_light;  /* BaseSeedLight, default 0x13C0FEFE */
    int32_t num_light_paths;   /* NumLightPaths, < 0 (default) = one per pixel */
    int32_t max_depth;         /* MaxDepth (Integrator default 100): at most max_depth - 1 light vertices */
    int32_t min_depth;         /* MinDepth (default 1): shorter full paths are not splatted */
} sspt_lt_settings;
sspt_status sspt_render_light_tracer(sspt_ctx* ctx, sspt_scene* scene, const sspt_lt_settings* settings,
                                     uint32_t first_iteration, uint32_t num_iterations,
                                     sspt_ray_stats* stats);

/* ---- bidirectional path tracing: ClassicBidir (SeeSharp/Integrators/Bidir/ClassicBidir.cs)
 *      over BidirBase.Render (BidirBase.cs:84-146). --------------------------
 * Each iteration first traces one light path per pixel (seeded
 * RNG(base_seed_light, path, iteration), as sspt_render_light_tracer) and keeps
 * its vertices (PathCache, LightPathWalk.OnHit / OnContinue,
 * BidirBase.Light.cs:324-410); if enable_light_tracer, every vertex is
 * connected to the camera (ConnectLightVertexToCamera, BidirBase.Light.cs:35-92).
 * Then one camera path per pixel (RNG(base_seed_camera, row*W+col, iteration),
 * BidirBase.RenderPixel, BidirBase.Camera.cs:97-106, a RandomWalk of
 * max_depth + 1 vertices) adds at every vertex the emitter it hit
 * (OnEmitterHit, :500-524), if enable_connections the connections to every
 * vertex of the light path of its pixel (BidirConnections / Connect,
 * :192-307), and num_shadow_rays next-event samples (PerformNextEventEstimation,
 * :368-484), all weighted by the balance heuristic over the enabled techniques
 * (ClassicBidir.EmitterHitMis / LightTracerMis / BidirConnectMis /
 * NextEventMis / CameraPathReciprocals / LightPathReciprocals,
 * ClassicBidir.cs:94-198, with the pdfs of BidirPathPdfs.cs). Splats of the light
 * tracer and the camera paths' pixel values go into the framebuffer with
 * running-mean semantics (as sspt_render_light_tracer: scale by (n-1)/n, then
 * add every finite value / n, the light tracer's in an unspecified order);
 * non-finite values are counted. The light-path cache holds every vertex (the
 * reference's PathCache of min(MaxDepth + 1, 10) vertices per path drops whole
 * paths only on overflow, which cannot happen for MaxDepth <= 9). With a
 * background (EnvironmentMap), SelectLight picks it with probability
 * 1 / (1 + emitters) and the light path starts from EnvironmentMap.SampleRay
 * (a map direction and a point on the disc of radius SceneRadius about the
 * scene's vertex mean, BidirBase.Light.cs:103-133,174-189,283-299,
 * RandomWalk.cs:89-128); next-event estimation picks it with the same
 * probability and camera paths that leave the scene add its radiance under
 * EmitterHitMis (BidirBase.Camera.cs:338-418,532-561), as the reference does
 * (no MinDepth test for these hits; the background's next-event pdf without
 * the NumShadowRays factor). Batches are SSPT_ERR_UNSUPPORTED;
 * num_light_paths other than -1 or the pixel count is SSPT_ERR_INVALID
 * (ClassicBidir.cs:72-75). */
typedef struct sspt_bidir_settings {
    uint32_t base_seed_camera;    /* BaseSeedCamera, default 0xC030114 */
    uint32_t base_seed_light;     /* BaseSeedLight, default 0x13C0FEFE */
    int32_t max_depth;            /* MaxDepth (Integrator default 100) */
    int32_t min_depth;            /* MinDepth (default 1) */
    int32_t num_shadow_rays;      /* NumShadowRays (default 1) */
    int32_t enable_connections;   /* EnableConnections (default 1) */
    int32_t enable_light_tracer;  /* EnableLightTracer (default 1) */
    int32_t num_light_paths;      /* NumLightPaths: -1 (default) or width * height */
} sspt_bidir_settings;
sspt_status sspt_render_bidir(sspt_ctx* ctx, sspt_scene* scene, const sspt_bidir_settings* settings,
                              uint32_t first_iteration, uint32_t num_iterations, sspt_ray_stats* stats);

/* ---- light vertex cache and merging: VertexCacheBidir
 *      (SeeSharp/Integrators/Bidir/VertexCacheBidir.cs) and VertexConnectionAndMerging
 *      (SeeSharp/Integrators/Bidir/VertexConnectionAndMerging.cs) over BidirBase.Render. --
 * Light paths and the light tracer as sspt_render_bidir, but num_light_paths of
 * them (< 0: one per pixel) and the light tracer weighted against the
 * techniques below. Each camera vertex below MaxDepth makes num_connections
 * connections, each to one vertex drawn uniformly (RNG.NextInt of the camera
 * path's generator) from all cached vertices in path order, light points
 * included, which connect to nothing (VertexSelector.cs,
 * BidirBase.Camera.cs:279-291); the connection's MIS density is
 * num_connections * num_light_paths / vertex count (BidirSelectDensity,
 * VertexCacheBidir.cs:71-82). enable_hitting switches the emitter hits of
 * camera paths (VCM keeps those of primary hits). SSPT_VCM adds merging
 * (PerformMerging / MergeHelper / Merge, VertexConnectionAndMerging.cs:329-516): at every
 * camera vertex past the primary hit (or at it too with merge_primary), the
 * max_num_photons light vertices (of depth >= 1 and non-zero weight) nearest
 * to it within the radius of the camera path's primary pixel footprint
 * (sqrt(1 / area pdf of the primary hit)) are merged with the Epanechnikov
 * kernel, whose radius is the distance of the furthest found photon when
 * max_num_photons were found. The k nearest are the stored points with the k
 * smallest squared distances (fp32, dx*dx + dy*dy + dz*dz) below the squared
 * radius, ties to the earlier photon: the query of TinyEmbree 1.2.0's
 * NearestNeighborSearch.ForAllNearest, restated (the package is not part of
 * the reference's sources). All techniques are combined by the balance
 * heuristic, for SSPT_VCM with the correlation-aware merge weights (Grittmann
 * et al. 2021, CorrelAwareRatios / MergeMis :580-697) unless disable_correl_aware_mis.
 * Backgrounds as for sspt_render_bidir (hits added whatever enable_hitting,
 * CorrelAwareRatios with fromBackground). Batches and max_num_photons outside
 * 1..32 are rejected. */
typedef enum sspt_vcm_algorithm {
    SSPT_VERTEX_CACHE_BIDIR = 0,  /* VertexCacheBidir */
    SSPT_VCM = 1                  /* VertexConnectionAndMerging */
} sspt_vcm_algorithm;
typedef struct sspt_vcm_settings {
    int32_t algorithm;                 /* sspt_vcm_algorithm */
    uint32_t base_seed_camera;         /* BaseSeedCamera, default 0xC030114 */
    uint32_t base_seed_light;          /* BaseSeedLight, default 0x13C0FEFE */
    int32_t max_depth;                 /* MaxDepth (Integrator default 100) */
    int32_t min_depth;                 /* MinDepth (default 1) */
    int32_t num_light_paths;           /* NumLightPaths, < 0 (default) = one per pixel */
    int32_t num_connections;           /* NumConnections (default 1) */
    int32_t num_shadow_rays;           /* NumShadowRays (default 1) */
    int32_t enable_light_tracer;       /* EnableLightTracer (default 1) */
    int32_t enable_hitting;            /* EnableHitting (default 1) */
    int32_t enable_merging;            /* VCM: EnableMerging (default 1) */
    int32_t merge_primary;             /* VCM: MergePrimary (default 0) */
    int32_t max_num_photons;           /* VCM: MaxNumPhotons (default 8) */
    int32_t disable_correl_aware_mis;  /* VCM: DisableCorrelAwareMIS (default 0) */
} sspt_vcm_settings;
sspt_status sspt_render_vcm(sspt_ctx* ctx, sspt_scene* scene, const sspt_vcm_settings* settings,
                            uint32_t first_iteration, uint32_t num_iterations, sspt_ray_stats* stats);

/* ---- measurement ------------------------------------------------------- */
/* Enables HIP-event timing of every kernel launch on the context stream. */
sspt_status sspt_set_kernel_timing(sspt_ctx* ctx, int32_t enable);
/* Accumulated launches and milliseconds of kernel `name` since the last reset:
 * "extend" (the paths' start and their camera rays; with the unfused schedule
 * "generate" and "extend"), "shade", "trace" (a bounce's
 * continuation rays with the previous bounce's shadow rays), "shadow" (the
 * last bounce's shadow rays), "accumulate", "path" (single-kernel tracer),
 * and "shade_adjoint", "shade_grad" (the adjoint's material derivatives),
 * "trace_adjoint", "shadow_adjoint", "adjoint_finalize"; "light_paths" (the
 * light tracer), "bidir_light" and "bidir_camera" (bidirectional rendering). */
sspt_status sspt_get_kernel_timing(sspt_ctx* ctx, const char* name, uint64_t* launches,
                                   double* total_ms);
sspt_status sspt_reset_kernel_timing(sspt_ctx* ctx);

#ifdef __cplusplus
}
#endif

#endif /* SSPT_H */
