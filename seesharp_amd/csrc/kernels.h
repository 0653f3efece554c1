// kernels.h -- launch interface of the wavefront path-tracer kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "scene_types.h"

namespace sspt {

// Wavefront buffers, all in HBM. P = path slots of one wave (iterations x pixels).
struct Wave {
    uint32_t capacity;     // P
    // extension queues (double buffered): origin+tmin, direction+tmax, path slot
    float4* q_o[2];
    float4* q_d[2];
    int* q_slot[2];
    float4* hits;          // (t, u, v, leaf triangle index bits), per queue entry
    // shadow queue: origin+tmin, direction+tmax, contribution rgb + (slot<<1 | kind)
    float4* s_o;
    float4* s_d;
    float4* s_c;
    // path state per slot
    uint32_t* st_rng;
    float4* st_thr;   // prefix weight rgb, previous pdf
    float4* st_prev;  // previous hit position, depth (int bits)
    float4* st_L;     // radiance from emitter / background hits
    float4* st_Le;    // environment next-event contributions
    float4* st_La;    // area-light next-event contributions
    uint32_t* counters;          // kCntNum
    unsigned long long* stats;   // kStatNum
};

void launch_generate(hipStream_t s, const DevScene& sc, const Wave& w, uint32_t base_seed,
                     uint32_t first_iter, uint32_t num_iters);
void launch_extend(hipStream_t s, const DevScene& sc, const Wave& w, int buf, uint32_t count);
void launch_shade(hipStream_t s, const DevScene& sc, const Wave& w, const DevSettings& cfg, int buf,
                  uint32_t count);
void launch_shadow(hipStream_t s, const DevScene& sc, const Wave& w, uint32_t count);
void launch_accumulate(hipStream_t s, const Wave& w, float* image, uint32_t num_pixels,
                       uint32_t first_iter, uint32_t num_iters);

// Ray queries for the API: rays are 8 floats (ox,oy,oz,tmin,dx,dy,dz,tmax).
void launch_trace_closest(hipStream_t s, const DevScene& sc, const float* rays, uint64_t n, float4* hits_out);
void launch_trace_occluded(hipStream_t s, const DevScene& sc, const float* rays, uint64_t n, uint8_t* occ);
void launch_primary_rays(hipStream_t s, const DevScene& sc, uint32_t base_seed, uint32_t iteration, float* rays);

}  // namespace sspt
