/*
 * oracle.h -- CPU restatement of SeeSharp's unidirectional PathTracer.
 *
 * TEST INFRASTRUCTURE ONLY. This library is the checker the GPU path is
 * compared against and the CPU baseline that bench.py times beside it. The
 * product (libsspt.so and the seesharp_amd package) never links, loads or
 * calls it. Each function cites the reference file:line it restates.
 */
#ifndef SSPT_ORACLE_H
#define SSPT_ORACLE_H

#include <stdint.h>

#include "../include/sspt.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---- RNG (SeeSharp/Sampling/RNG.cs) ---- */
uint32_t or_hash_seed(uint32_t base_seed, uint32_t chain_index, uint32_t sample_index);
uint32_t or_pcg_hash(uint32_t x);
uint32_t or_fnv_hash(uint32_t hash, uint32_t data);
/* n successive Next() values and NextFloat() values of one generator. */
void or_rng_stream(uint32_t seed, uint32_t n, uint32_t* out_u32, float* out_float);
/* NextInt(max) n times from one generator. */
void or_rng_next_int(uint32_t seed, uint32_t max, uint32_t n, uint32_t* out);

/* ---- SampleWarp (SeeSharp/Sampling/SampleWarp.cs) ---- */
void or_warp_uniform_triangle(const float* primary2, float* out_bary2);
void or_warp_cos_hemisphere(const float* primary2, float* out_dir3, float* out_pdf);

/* ---- PiecewiseConstantPDF (SeeSharp/Sampling/PiecewiseConstant.cs) ---- */
void or_pwc_sample(const float* weights, int32_t n, float u, int32_t* out_idx, float* out_rel);
float or_pwc_probability(const float* weights, int32_t n, int32_t idx);

/* ---- camera (SeeSharp/Cameras/PerspectiveCamera.cs) ---- */
/* ray_out: ox,oy,oz,dx,dy,dz */
void or_camera_ray(const float* world_to_camera16, float fov_deg, int32_t width, int32_t height,
                   float film_x, float film_y, float* ray_out);
void or_look_at(const float* pos3, const float* target3, const float* up3, float* out16);

/* ---- microfacet + fresnel (SeeSharp/Shading/TrowbridgeReitzDistribution.cs, Fresnel.cs) ---- */
float or_ggx_ndf(float ax, float ay, const float* n3);
float or_ggx_pdf(float ax, float ay, const float* out3, const float* n3);
void or_ggx_sample(float ax, float ay, const float* out3, const float* primary2, float* out_n3);
float or_fresnel_dielectric(float cos_i, float eta_i, float eta_t);

/* ---- materials on a flat test surface (SeeSharp/Shading/Materials/) ----
 * The surface point has face normal = shading normal = normal3, texture
 * coordinates uv2, and the material's textures taken from `textures`. */
void or_material_eval(const sspt_material_desc* mat, const sspt_texture_desc* textures,
                      const float* normal3, const float* uv2, const float* out_dir3,
                      const float* in_dir3, float* value3, float* value_cos3, float* pdf);
/* SurfaceShader.Pdf(inDir) (SurfaceShader.cs:64-67): the forward and the reverse pdf. */
void or_material_pdf_pair(const sspt_material_desc* mat, const sspt_texture_desc* textures, const float* normal3,
                          const float* uv2, const float* out_dir3, const float* in_dir3, float* pdf,
                          float* pdf_rev);
/* f * |cos| (3 floats each) and the forward pdf (may be NULL) of n incident
 * directions in_dirs (3 floats each) for one outgoing direction. */
void or_material_eval_many(const sspt_material_desc* mat, const sspt_texture_desc* textures,
                           const float* normal3, const float* uv2, const float* out_dir3, int64_t n,
                           const float* in_dirs, float* value_cos, float* pdf);
/* sample_out: dir3, pdf, weight3, reverse pdf -> 8 floats */
void or_material_sample(const sspt_material_desc* mat, const sspt_texture_desc* textures,
                        const float* normal3, const float* uv2, const float* out_dir3,
                        float primary_component, const float* primary2, float* sample_out);

/* Monte Carlo integral over n samples: mode 0 pdf (uniform sphere), 1 f*|cos|
 * (uniform sphere), 2 BSDF-sample weight. Returns the mean and the variance. */
void or_material_integrate(const sspt_material_desc* mat, const sspt_texture_desc* textures,
                           const float* normal3, const float* uv2, const float* out_dir3, int64_t n,
                           uint32_t seed, int32_t mode, double* mean, double* var);

/* ---- scene (SeeSharp/Scene.cs Prepare) ---- */
typedef struct or_scene or_scene;
/* Returns NULL and writes a message to err (if non-NULL, 256 bytes) on failure. */
or_scene* or_scene_create(const sspt_scene_desc* desc, char* err);
void or_scene_destroy(or_scene* scene);

/* EnvironmentMap of the scene (SeeSharp/Shading/Background/EnvironmentMap.cs) */
void or_env_eval(const or_scene* s, const float* dir3, float* rgb3);
float or_env_pdf(const or_scene* s, const float* dir3);
/* out: dir3, pdf, weight3 */
void or_env_sample(const or_scene* s, const float* primary2, float* out7);

/* ---- ray queries: rays are 8 floats (ox,oy,oz,tmin,dx,dy,dz,tmax) ---- */
void or_trace_closest(const or_scene* s, const float* rays, uint64_t n, sspt_hit* hits,
                      int32_t brute_force);
void or_trace_occluded(const or_scene* s, const float* rays, uint64_t n, uint8_t* occluded,
                       int32_t brute_force);
void or_primary_hits(const or_scene* s, uint32_t base_seed, uint32_t iteration,
                     float* rays_out, sspt_hit* hits_out, int32_t brute_force);

/* ---- PathTracerBase.Render (SeeSharp/Integrators/PathTracer.cs:199-382) ----
 * Renders iterations first..first+num-1 into image (W*H*3 running mean) for
 * all pixels, or only the `num_pixels` pixel indices listed in `pixels`.
 * num_threads <= 0 uses the hardware threads, at most 16. Returns the NaN/Inf splat count. */
uint64_t or_render(const or_scene* s, const sspt_pt_settings* settings, uint32_t first_iteration,
                   uint32_t num_iterations, float* image, sspt_ray_stats* stats,
                   const int32_t* pixels, int64_t num_pixels, int32_t num_threads);
/* The path estimate of one (pixel, iteration), PathTracer.RenderPixel without the splat. */
void or_render_pixel(const or_scene* s, const sspt_pt_settings* settings, uint32_t row,
                     uint32_t col, uint32_t iteration, float* rgb3);

/* DenoiseBuffers layers: running means of the primary hit's GetScatterStrength
 * (albedo) and shading normal over iterations first..first+num-1, W*H*3 each,
 * updated in place (start from zeros for a fresh framebuffer). */
/* The FrameBuffer's outlier cache after rendering iterations first..first+num-1:
 * per pixel the n largest samples by average estimate, W*H*n weights and
 * iterations in descending weight (weight 0 / iteration -1 past the kept count). */
void or_render_outliers(const or_scene* s, const sspt_pt_settings* settings, uint32_t first, uint32_t num, int32_t n,
                        float* weights, int32_t* iterations, int32_t threads);
/* DebugVisualizer: the eye-light image of iterations first..first+num-1 into the
 * running-mean W*H*3 image. */
void or_render_debug(const or_scene* s, uint32_t base_seed, uint32_t first, uint32_t num, float* image,
                     int32_t threads);
/* PathTracerBase.ReplayPixel: the path graph (sspt_path_node array, at most
 * `capacity` written) and estimate of pixel (col, row) in `iteration`;
 * returns the node count. */
uint32_t or_replay_pixel(const or_scene* s, const sspt_pt_settings* settings, uint32_t col, uint32_t row,
                         uint32_t iteration, sspt_path_node* nodes, uint32_t capacity, float* estimate_rgb);
/* ---- light tracing: BidirBase.Render (SeeSharp/Integrators/Bidir/BidirBase.cs:84-146,
 * BidirBase.Light.cs:35-277, RandomWalk.cs) of a VertexCacheBidir set up as a pure
 * light tracer, the contract of sspt_render_light_tracer; the splats of an
 * iteration are added in light-path order. Returns the non-finite splat count,
 * or UINT64_MAX for a scene with a background or without emitters. */
uint64_t or_render_light_tracer(const or_scene* s, const sspt_lt_settings* settings, uint32_t first_iteration,
                                uint32_t num_iterations, float* image, sspt_ray_stats* stats,
                                int32_t num_threads);
/* ---- bidirectional path tracing: ClassicBidir (SeeSharp/Integrators/Bidir/ClassicBidir.cs),
 * the contract of sspt_render_bidir; per iteration the light tracer's splats
 * are added in light-path order, then the camera paths' values in pixel
 * order. Returns the non-finite value count, UINT64_MAX for a scene with a
 * background or without emitters or for num_light_paths other than -1 or
 * the pixel count. */
uint64_t or_render_bidir(const or_scene* s, const sspt_bidir_settings* settings, uint32_t first_iteration,
                         uint32_t num_iterations, float* image, sspt_ray_stats* stats, int32_t num_threads);
/* ---- VertexCacheBidir / VertexConnectionAndMerging (SeeSharp/Integrators/Bidir/
 * VertexCacheBidir.cs, VertexConnectionAndMerging.cs), the contract of
 * sspt_render_vcm; the photon map is a uniform grid of cells of cell_size
 * (any size gives the same photons, it only changes the work). Splats in
 * light-path order, then the camera paths' values in pixel order. Returns the
 * non-finite value count, UINT64_MAX for a scene with a background or without
 * emitters, or invalid settings. */
uint64_t or_render_vcm(const or_scene* s, const sspt_vcm_settings* settings, float cell_size,
                       uint32_t first_iteration, uint32_t num_iterations, float* image, sspt_ray_stats* stats,
                       int32_t num_threads);
void or_primary_aov(const or_scene* s, const sspt_pt_settings* settings, uint32_t first_iteration,
                    uint32_t num_iterations, float* albedo, float* normal, int32_t num_threads);

/* ---- derivative reference for the path-replay adjoint ----------------------
 * Replaces the texels of texture `index` (scene-description order). Returns 0,
 * or 1 for an index out of range. */
int32_t or_set_texture(or_scene* s, uint32_t index, const float* texels);
/* Directional derivative of each pixel's estimate (mean over iterations
 * first..first+num-1) when texture `texture_index` moves along `tangent`:
 * every path is traced once with the scene's texels, which fixes all samples,
 * pdfs, MIS weights and Russian-roulette decisions, and its estimate is also
 * accumulated with the BSDF values of texels +- eps * tangent; the derivative
 * is their central difference. This is the quantity the adjoint
 * differentiates (sampling held fixed). Writes W*H*3 floats at the listed
 * pixels (all pixels if `pixels` is null); returns the skipped non-finite
 * estimates. */
uint64_t or_render_tangent(const or_scene* s, const sspt_pt_settings* settings, uint32_t first_iteration,
                           uint32_t num_iterations, uint32_t texture_index, const float* tangent,
                           float eps, float* dimage, const int32_t* pixels, int64_t num_pixels,
                           int32_t num_threads);

#ifdef __cplusplus
}
#endif

#endif
