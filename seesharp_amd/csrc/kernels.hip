// kernels.hip -- the wavefront path tracer for gfx950.
//
// One iteration batch (several 1-spp iterations of every pixel) is a wave of
// path slots. Each bounce runs
//   extend  : closest-hit BVH traversal of the compacted extension queue
//   shade   : emitter/background hits with MIS, Russian roulette, next-event
//             sampling (environment + area light) and BSDF sampling; appends
//             shadow rays and continuation rays with wave64 ballot compaction
//   shadow  : any-hit traversal; visible contributions go to the path slot
// and the batch ends with `accumulate`, the FrameBuffer running mean.
// This is the GPU form of PathTracerBase.Render / RenderPixel /
// EstimateIncidentRadiance (SeeSharp/Integrators/PathTracer.cs:199-530).
#include "kernels.h"
#include "shading.h"

namespace sspt {

// ------------------------------------------------------------ wave helpers -
// Appends one entry per lane with `pred` set: one atomic per wave64.
__device__ __forceinline__ uint32_t wave_append(uint32_t* counter, bool pred) {
    uint64_t mask = __ballot(pred);
    if (mask == 0) return 0;
    uint32_t lane = __lane_id();
    int leader = __ffsll((unsigned long long)mask) - 1;
    uint32_t base = 0;
    if ((int)lane == leader) base = atomicAdd(counter, (uint32_t)__popcll(mask));
    base = __shfl(base, leader);
    return base + (uint32_t)__popcll(mask & ((1ull << lane) - 1ull));
}
__device__ __forceinline__ void wave_count(unsigned long long* stat, bool pred) {
    uint64_t mask = __ballot(pred);
    if (mask == 0) return;
    int leader = __ffsll((unsigned long long)mask) - 1;
    if ((int)__lane_id() == leader) atomicAdd(stat, (unsigned long long)__popcll(mask));
}

// ------------------------------------------------------------------ camera -
// PerspectiveCamera.GenerateRay (SeeSharp/Cameras/PerspectiveCamera.cs:64-102), lensRadius = 0:
// view = (2x/W - 1, 1 - 2y/H, 0); local = Transform(view, viewToCamera) (w = 1);
// dir = normalize(Transform((local, 0), cameraToWorld)); origin = camera position.
__device__ __forceinline__ void camera_ray(const DevCamera& c, float fx, float fy, f3* o, f3* d) {
    float vx = 2.0f * fx / (float)c.width - 1.0f;
    float vy = 1.0f - 2.0f * fy / (float)c.height;
    float vz = 0.0f;
    const float* M = c.v2c;
    f3 l = mk3(vx * M[0] + vy * M[4] + vz * M[8] + M[12], vx * M[1] + vy * M[5] + vz * M[9] + M[13],
               vx * M[2] + vy * M[6] + vz * M[10] + M[14]);
    const float* W = c.c2w;
    f3 dd = mk3(l.x * W[0] + l.y * W[4] + l.z * W[8], l.x * W[1] + l.y * W[5] + l.z * W[9],
                l.x * W[2] + l.y * W[6] + l.z * W[10]);
    *o = mk3(c.pos[0], c.pos[1], c.pos[2]);
    *d = normalize(dd);
}

// ------------------------------------------------------------ intersection -
struct RayPre {
    int kx, ky, kz;
    float sx, sy, sz;
    f3 inv;
};
__device__ __forceinline__ RayPre precompute(f3 d) {
    RayPre p;
    float ax = fabsf(d.x), ay = fabsf(d.y), az = fabsf(d.z);
    p.kz = (ax > ay) ? (ax > az ? 0 : 2) : (ay > az ? 1 : 2);
    p.kx = p.kz == 2 ? 0 : p.kz + 1;
    p.ky = p.kx == 2 ? 0 : p.kx + 1;
    if (comp(d, p.kz) < 0.0f) {
        int t = p.kx;
        p.kx = p.ky;
        p.ky = t;
    }
    float dz = comp(d, p.kz);
    p.sx = comp(d, p.kx) / dz;
    p.sy = comp(d, p.ky) / dz;
    p.sz = 1.0f / dz;
    p.inv = mk3(1.0f / d.x, 1.0f / d.y, 1.0f / d.z);
    return p;
}

// Watertight ray/triangle test (Woop, Benthin, Wald, JCGT 2(1) 2013) with the
// double-precision fallback for zero edge functions. u, v weight vertices 1, 2.
__device__ __forceinline__ bool intersect_tri(f3 o, const RayPre& p, float tmin, float tmax, f3 p0,
                                              f3 p1, f3 p2, float* t_out, float* u_out, float* v_out) {
    f3 a = p0 - o, b = p1 - o, c = p2 - o;
    float az_ = comp(a, p.kz), bz_ = comp(b, p.kz), cz_ = comp(c, p.kz);
    float ax = comp(a, p.kx) - p.sx * az_;
    float ay = comp(a, p.ky) - p.sy * az_;
    float bx = comp(b, p.kx) - p.sx * bz_;
    float by = comp(b, p.ky) - p.sy * bz_;
    float cx = comp(c, p.kx) - p.sx * cz_;
    float cy = comp(c, p.ky) - p.sy * cz_;
    float U = cx * by - cy * bx;
    float V = ax * cy - ay * cx;
    float W = bx * ay - by * ax;
    if (U == 0.0f || V == 0.0f || W == 0.0f) {
        double cxby = (double)cx * (double)by, cybx = (double)cy * (double)bx;
        U = (float)(cxby - cybx);
        double axcy = (double)ax * (double)cy, aycx = (double)ay * (double)cx;
        V = (float)(axcy - aycx);
        double bxay = (double)bx * (double)ay, byax = (double)by * (double)ax;
        W = (float)(bxay - byax);
    }
    if ((U < 0.0f || V < 0.0f || W < 0.0f) && (U > 0.0f || V > 0.0f || W > 0.0f)) return false;
    float det = U + V + W;
    if (det == 0.0f) return false;
    float T = U * (p.sz * az_) + V * (p.sz * bz_) + W * (p.sz * cz_);
    float rcp = 1.0f / det;
    float t = T * rcp;
    if (!(t >= tmin && t <= tmax)) return false;
    *t_out = t;
    *u_out = V * rcp;
    *v_out = W * rcp;
    return true;
}

// Robust slab test (Ize, JCGT 2013): far distances widened by 1 + 2 gamma(3).
// The caller widens the query interval by the ray's slack (see traverse).
__device__ __forceinline__ bool box_test(float lox, float hix, float loy, float hiy, float loz, float hiz,
                                         f3 o, f3 inv, float tmin, float tmax, float* tnear) {
    float x0 = (lox - o.x) * inv.x, x1 = (hix - o.x) * inv.x;
    float y0 = (loy - o.y) * inv.y, y1 = (hiy - o.y) * inv.y;
    float z0 = (loz - o.z) * inv.z, z1 = (hiz - o.z) * inv.z;
    float tn = fmaxf(fmaxf(fminf(x0, x1), fminf(y0, y1)), fmaxf(fminf(z0, z1), tmin));
    float tf = fminf(fminf(fmaxf(x0, x1) * 1.0000004f, fmaxf(y0, y1) * 1.0000004f),
                     fminf(fmaxf(z0, z1) * 1.0000004f, tmax));
    *tnear = tn;
    return tn <= tf;
}

// The watertight t carries an absolute error of a few ulps of the coordinate
// magnitudes (divided by |d|), not of t itself. Box culling against the best
// hit so far uses [tmin - slack, tmax + slack] with slack = 2^-16 * max(|o|,
// scene extent) / |d|, so a box is never culled by rounding against a triangle
// distance that is equal in exact arithmetic.
__device__ __forceinline__ float ray_slack(const DevScene& S, f3 o, f3 d) {
    float m = fmaxf(S.abs_max, fmaxf(fmaxf(fabsf(o.x), fabsf(o.y)), fabsf(o.z)));
    return m * 1.52587890625e-05f / length(d);
}

struct HitOut {
    float t, u, v;
    int li;   // leaf-order triangle index, -1 = miss
    int gid;  // global triangle id
};

// BVH traversal with a per-lane stack in LDS (stack[depth][lane] layout: no
// bank conflicts). ANY = true stops at the first hit (shadow rays).
template <bool ANY>
__device__ __forceinline__ HitOut traverse(const DevScene& S, f3 o, f3 d, float tmin, float tmax, int* stk) {
    RayPre p = precompute(d);
    const float slack = ray_slack(S, o, d);
    const float tmin_box = tmin - slack;
    HitOut best{tmax, 0.0f, 0.0f, -1, 0x7fffffff};
    int sp = 0;
    int node = 0;
    const int stride = kTraceBlock;
    for (;;) {
        const float4* N = S.nodes + 4 * node;
        float4 n0 = N[0], n1 = N[1], n2 = N[2];
        int4 c = *reinterpret_cast<const int4*>(N + 3);
        float tl, tr;
        const float tmax_box = best.t + slack;
        bool hl = box_test(n0.x, n0.y, n0.z, n0.w, n2.x, n2.y, o, p.inv, tmin_box, tmax_box, &tl);
        bool hr = box_test(n1.x, n1.y, n1.z, n1.w, n2.z, n2.w, o, p.inv, tmin_box, tmax_box, &tr);
        // leaf children are intersected right away
        for (int side = 0; side < 2; ++side) {
            bool h = side == 0 ? hl : hr;
            int cnt = side == 0 ? c.z : c.w;
            if (!(h && cnt > 0)) continue;
            int first = side == 0 ? c.x : c.y;
            for (int k = first; k < first + cnt; ++k) {
                float4 a = S.tri_pos[3 * k], b = S.tri_pos[3 * k + 1], cc = S.tri_pos[3 * k + 2];
                float t, u, v;
                if (!intersect_tri(o, p, tmin, best.t, f3_of(a), f3_of(b), f3_of(cc), &t, &u, &v)) continue;
                int gid = __float_as_int(a.w);
                if (ANY) {
                    best.t = t;
                    best.u = u;
                    best.v = v;
                    best.li = k;
                    best.gid = gid;
                    return best;
                }
                if (t < best.t || gid < best.gid) {
                    best.t = t;
                    best.u = u;
                    best.v = v;
                    best.li = k;
                    best.gid = gid;
                }
            }
            if (side == 0) hl = false;
            else hr = false;
        }
        if (hl && hr) {
            int near = tl <= tr ? c.x : c.y;
            int far = tl <= tr ? c.y : c.x;
            stk[sp * stride] = far;
            ++sp;
            node = near;
        } else if (hl) {
            node = c.x;
        } else if (hr) {
            node = c.y;
        } else {
            if (sp == 0) break;
            --sp;
            node = stk[sp * stride];
        }
    }
    return best;
}

// ---------------------------------------------------------------- kernels --
__global__ void __launch_bounds__(256) k_generate(DevScene S, Wave w, uint32_t base_seed,
                                                  uint32_t first_iter, uint32_t num_pixels,
                                                  uint32_t total) {
    uint32_t slot = blockIdx.x * blockDim.x + threadIdx.x;
    if (slot >= total) return;
    uint32_t it = slot / num_pixels;
    uint32_t pix = slot - it * num_pixels;
    uint32_t row = pix / (uint32_t)S.width;
    uint32_t col = pix - row * (uint32_t)S.width;
    // RNG(BaseSeed, row * W + col, sampleIndex) (PathTracer.cs:236-237); jitter (:289-290)
    Rng rng{hash_seed(base_seed, pix, first_iter + it)};
    f2 off = rng.next_float2();
    f3 o, d;
    camera_ray(S.cam, (float)col + off.x, (float)row + off.y, &o, &d);
    w.q_o[0][slot] = f4_of(o, 0.0f);
    w.q_d[0][slot] = f4_of(d, INFINITY);
    w.q_slot[0][slot] = (int)slot;
    w.st_rng[slot] = rng.state;
    w.st_thr[slot] = make_float4(1.0f, 1.0f, 1.0f, 0.0f);
    w.st_prev[slot] = make_float4(0.0f, 0.0f, 0.0f, __int_as_float(1));
    w.st_L[slot] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    w.st_Le[slot] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    w.st_La[slot] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
}

__global__ void __launch_bounds__(kTraceBlock) k_extend(DevScene S, const float4* __restrict__ q_o,
                                                        const float4* __restrict__ q_d, float4* hits,
                                                        uint32_t count, unsigned long long* stats) {
    __shared__ int stack[kMaxBvhDepth * kTraceBlock];
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    float4 o4 = q_o[i], d4 = q_d[i];
    HitOut h = traverse<false>(S, f3_of(o4), f3_of(d4), o4.w, d4.w, stack + threadIdx.x);
    hits[i] = h.li >= 0 ? make_float4(h.t, h.u, h.v, __int_as_float(h.li))
                        : make_float4(INFINITY, 0.0f, 0.0f, __int_as_float(-1));
    wave_count(&stats[kStatRays], true);
    wave_count(&stats[kStatHits], h.li >= 0);
}

__global__ void __launch_bounds__(kTraceBlock) k_shadow(DevScene S, const float4* __restrict__ s_o,
                                                        const float4* __restrict__ s_d,
                                                        const float4* __restrict__ s_c, float4* L_env,
                                                        float4* L_area, uint32_t count,
                                                        unsigned long long* stats) {
    __shared__ int stack[kMaxBvhDepth * kTraceBlock];
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    float4 o4 = s_o[i], d4 = s_d[i];
    HitOut h = traverse<true>(S, f3_of(o4), f3_of(d4), o4.w, d4.w, stack + threadIdx.x);
    bool occluded = h.li >= 0;
    if (!occluded) {
        float4 c = s_c[i];
        int tag = __float_as_int(c.w);
        float4* dst = (tag & 1) ? L_area : L_env;
        int slot = tag >> 1;
        atomicAdd(&dst[slot].x, c.x);
        atomicAdd(&dst[slot].y, c.y);
        atomicAdd(&dst[slot].z, c.z);
    }
    wave_count(&stats[kStatShadow], true);
    wave_count(&stats[kStatOccluded], occluded);
}

// TinyEmbree SpawnRay / IsOccluded offset rule: move along the face normal by
// the error offset, to the side the ray leaves on.
__device__ __forceinline__ f3 offset_point(const Surf& s, f3 dir) {
    float sign = dot(dir, s.ng) < 0.0f ? -1.0f : 1.0f;
    return s.pos + sign * s.err * s.ng;
}

__global__ void __launch_bounds__(256) k_shade(DevScene S, Wave w, DevSettings cfg, int buf, uint32_t count) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    const int nb = buf ^ 1;
    int slot = w.q_slot[buf][i];
    f3 dir = f3_of(w.q_d[buf][i]);
    float4 h = w.hits[i];
    Rng rng{w.st_rng[slot]};
    float4 thr = w.st_thr[slot];
    f3 prefix = f3_of(thr);
    float prev_pdf = thr.w;
    float4 pv = w.st_prev[slot];
    int depth = __float_as_int(pv.w);
    f3 prev_pos = f3_of(pv);
    float4 L4 = w.st_L[slot];
    f3 L = f3_of(L4);
    const float nsr = (float)cfg.num_shadow_rays;
    const int li = __float_as_int(h.w);

    bool cont = false;
    bool do_nee = false;
    f3 new_dir = gray(0.0f);
    Surf s;
    Shader sh;
    float q = 1.0f;
    if (li < 0) {
        // OnBackgroundHit (PathTracer.cs:384-400)
        if (depth >= cfg.min_depth) {
            float mis = 0.0f;
            f3 contrib = gray(0.0f);
            if (S.has_env && cfg.enable_bsdf_di) {
                mis = 1.0f;
                if (depth > 1) {
                    float pdf_ne = env_pdf(S, dir) * nsr;
                    mis = 1.0f / (1.0f + pdf_ne / prev_pdf);
                }
                contrib = env_radiance(S, dir);
            }
            L = L + prefix * mis * contrib;
        }
    } else {
        s = surface(S, li, h.y, h.z);
        make_shader(S, s, -dir, &sh);
        int ei = S.tri_emitter[li];
        if (ei >= 0 && depth >= cfg.min_depth) {
            // OnLightHit (PathTracer.cs:402-421)
            DevEmitter e = S.emitters[ei];
            float mis = 1.0f;
            if (depth > 1) {
                float jac = area_to_solid_angle(prev_pos, s.pos, s.ng);
                float pdf_ne = e.inv_area / (float)S.num_emitters * nsr / jac;
                float ratio = pdf_ne / prev_pdf;
                mis = 1.0f / (ratio + 1.0f);
                if (!cfg.enable_bsdf_di) mis = 0.0f;
            }
            // DiffuseEmitter.EmittedRadiance: one-sided on the shading normal
            f3 em = dot(s.ns, -dir) < 0.0f ? gray(0.0f) : mk3(e.r, e.g, e.b);
            L = L + prefix * mis * em;
        }
        // ComputeSurvivalProbability (PathTracer.cs:277-282); ApproxThroughput == PrefixWeight here
        q = depth > 4 ? clampf(average(prefix), 0.05f, 0.95f) : 1.0f;
        bool stop = rng.next_float() > q || depth == cfg.max_depth;
        do_nee = !stop && depth + 1 >= cfg.min_depth;
        cont = !stop;
    }

    // Next-event estimation (PathTracer.cs:356-364): per shadow ray, environment then area light.
    for (int k = 0; k < cfg.num_shadow_rays; ++k) {
        // PerformBackgroundNextEvent (PathTracer.cs:423-452)
        bool env_ray = false;
        float4 eo, ed, ec;
        if (do_nee && S.has_env) {
            EnvSample es = env_sample(S, rng.next_float2());
            f3 bc;
            float pdf_bsdf;
            eval_cos_pdf(sh, es.dir, &bc, &pdf_bsdf);
            if (!(pdf_bsdf == 0.0f || es.pdf == 0.0f)) {
                float mis = cfg.enable_bsdf_di ? 1.0f / (1.0f + pdf_bsdf / (es.pdf * nsr)) : 1.0f;
                f3 contrib = es.weight * bc / nsr;
                f3 val = prefix * (mis * contrib) / q;
                env_ray = true;
                eo = f4_of(offset_point(s, es.dir), s.err);
                ed = f4_of(es.dir, INFINITY);
                ec = f4_of(val, __int_as_float(slot << 1));
            } else {
                env_ray = true;  // LeavesScene is still traced, contributes zero
                eo = f4_of(offset_point(s, es.dir), s.err);
                ed = f4_of(es.dir, INFINITY);
                ec = make_float4(0.0f, 0.0f, 0.0f, __int_as_float(slot << 1));
            }
        }
        uint32_t e_idx = wave_append(&w.counters[kCntShadow], env_ray);
        if (env_ray) {
            w.s_o[e_idx] = eo;
            w.s_d[e_idx] = ed;
            w.s_c[e_idx] = ec;
        }
        // PerformNextEventEstimation (PathTracer.cs:454-501)
        bool area_ray = false;
        float4 ao, ad, ac;
        if (do_nee && S.num_emitters > 0) {
            int idx = (int)rng.next_int((uint32_t)S.num_emitters);
            DevEmitter e = S.emitters[idx];
            float sel = 1.0f / (float)S.num_emitters;
            f2 b = to_uniform_triangle(rng.next_float2());
            Surf lp = surface(S, e.tri, b.x, b.y);
            f3 l2s = normalize(s.pos - lp.pos);
            f3 em = dot(lp.ns, l2s) < 0.0f ? gray(0.0f) : mk3(e.r, e.g, e.b);
            float jac = area_to_solid_angle(s.pos, lp.pos, lp.ng);
            f3 bc;
            float pdf_bsdf_sa;
            eval_cos_pdf(sh, -l2s, &bc, &pdf_bsdf_sa);
            float pdf_ne = e.inv_area * sel * nsr;
            float pdf_bsdf = pdf_bsdf_sa * jac;
            f3 val = gray(0.0f);
            if (jac != 0.0f) {
                float ratio = pdf_bsdf / pdf_ne;
                float mis = cfg.enable_bsdf_di ? 1.0f / (ratio + 1.0f) : 1.0f;
                float pdf = e.inv_area / jac * sel * nsr;
                f3 contrib = em / pdf * bc;
                val = prefix * (mis * contrib) / q;
            }
            f3 dd = lp.pos - s.pos;
            f3 p0 = offset_point(s, dd);
            f3 p1 = offset_point(lp, -dd);
            area_ray = true;
            ao = f4_of(p0, 1e-5f);
            ad = f4_of(p1 - p0, 1.0f - 1e-5f);
            ac = f4_of(val, __int_as_float((slot << 1) | 1));
        }
        uint32_t a_idx = wave_append(&w.counters[kCntShadow], area_ray);
        if (area_ray) {
            w.s_o[a_idx] = ao;
            w.s_d[a_idx] = ad;
            w.s_c[a_idx] = ac;
        }
    }

    // SampleDirection (PathTracer.cs:526-530) and the path-state update (:367-378)
    float4 no, nd;
    if (cont) {
        float u_comp = rng.next_float();
        f2 u_dir = rng.next_float2();
        BsdfSample bs = sample_bsdf(sh, u_comp, u_dir);
        if (bs.pdf == 0.0f || is_zero(bs.weight)) {
            cont = false;
        } else {
            new_dir = bs.dir;
            float sign = dot(bs.dir, s.ng) < 0.0f ? -1.0f : 1.0f;
            no = f4_of(s.pos + sign * s.err * s.ng, s.err);
            nd = f4_of(bs.dir, INFINITY);
            prefix = prefix * (bs.weight / q);
            w.st_thr[slot] = f4_of(prefix, bs.pdf * q);
            w.st_prev[slot] = f4_of(s.pos, __int_as_float(depth + 1));
        }
    }
    uint32_t n_idx = wave_append(&w.counters[kCntNext], cont);
    if (cont) {
        w.q_o[nb][n_idx] = no;
        w.q_d[nb][n_idx] = nd;
        w.q_slot[nb][n_idx] = slot;
    }
    w.st_rng[slot] = rng.state;
    w.st_L[slot] = f4_of(L, 0.0f);
    (void)new_dir;
}

// FrameBuffer running mean (SeeSharp/Images/FrameBuffer.cs:89-160): at
// iteration n the image is scaled by (n-1)/n and the finite estimate / n added.
__global__ void __launch_bounds__(256) k_accumulate(Wave w, float* image, uint32_t num_pixels,
                                                    uint32_t first_iter, uint32_t num_iters) {
    uint32_t pix = blockIdx.x * blockDim.x + threadIdx.x;
    if (pix >= num_pixels) return;
    f3 img = mk3(image[3 * pix], image[3 * pix + 1], image[3 * pix + 2]);
    uint32_t nans = 0;
    for (uint32_t s = 0; s < num_iters; ++s) {
        uint32_t slot = s * num_pixels + pix;
        f3 L = f3_of(w.st_L[slot]) + f3_of(w.st_Le[slot]) + f3_of(w.st_La[slot]);
        float n = (float)(first_iter + s + 1);
        if (first_iter + s + 1 > 1) img = img * ((n - 1.0f) / n);
        if (isfinite(average(L))) img = img + L / n;
        else ++nans;
    }
    image[3 * pix] = img.x;
    image[3 * pix + 1] = img.y;
    image[3 * pix + 2] = img.z;
    if (nans) atomicAdd(&w.stats[kStatNan], (unsigned long long)nans);
}

__global__ void __launch_bounds__(kTraceBlock) k_trace_closest(DevScene S, const float* rays, uint64_t n,
                                                               float4* out) {
    __shared__ int stack[kMaxBvhDepth * kTraceBlock];
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float* r = rays + i * 8;
    HitOut h = traverse<false>(S, mk3(r[0], r[1], r[2]), mk3(r[4], r[5], r[6]), r[3], r[7], stack + threadIdx.x);
    out[i] = h.li >= 0 ? make_float4(h.t, h.u, h.v, __int_as_float(h.gid))
                       : make_float4(INFINITY, 0.0f, 0.0f, __int_as_float(-1));
}

__global__ void __launch_bounds__(kTraceBlock) k_trace_occluded(DevScene S, const float* rays, uint64_t n,
                                                                uint8_t* out) {
    __shared__ int stack[kMaxBvhDepth * kTraceBlock];
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float* r = rays + i * 8;
    HitOut h = traverse<true>(S, mk3(r[0], r[1], r[2]), mk3(r[4], r[5], r[6]), r[3], r[7], stack + threadIdx.x);
    out[i] = h.li >= 0 ? 1 : 0;
}

__global__ void __launch_bounds__(256) k_primary_rays(DevScene S, uint32_t base_seed, uint32_t iteration,
                                                      float* rays) {
    uint32_t pix = blockIdx.x * blockDim.x + threadIdx.x;
    if (pix >= (uint32_t)(S.width * S.height)) return;
    uint32_t row = pix / (uint32_t)S.width, col = pix - row * (uint32_t)S.width;
    Rng rng{hash_seed(base_seed, pix, iteration)};
    f2 off = rng.next_float2();
    f3 o, d;
    camera_ray(S.cam, (float)col + off.x, (float)row + off.y, &o, &d);
    float* r = rays + (uint64_t)pix * 8;
    r[0] = o.x, r[1] = o.y, r[2] = o.z, r[3] = 0.0f;
    r[4] = d.x, r[5] = d.y, r[6] = d.z, r[7] = INFINITY;
}

// ------------------------------------------------------------------ launch --
static inline uint32_t blocks(uint64_t n, uint32_t b) { return (uint32_t)((n + b - 1) / b); }

void launch_generate(hipStream_t s, const DevScene& sc, const Wave& w, uint32_t base_seed, uint32_t first_iter,
                     uint32_t num_iters) {
    uint32_t np = (uint32_t)(sc.width * sc.height);
    uint32_t total = np * num_iters;
    hipLaunchKernelGGL(k_generate, dim3(blocks(total, 256)), dim3(256), 0, s, sc, w, base_seed, first_iter, np,
                       total);
}
void launch_extend(hipStream_t s, const DevScene& sc, const Wave& w, int buf, uint32_t count) {
    if (!count) return;
    hipLaunchKernelGGL(k_extend, dim3(blocks(count, kTraceBlock)), dim3(kTraceBlock), 0, s, sc, w.q_o[buf],
                       w.q_d[buf], w.hits, count, w.stats);
}
void launch_shade(hipStream_t s, const DevScene& sc, const Wave& w, const DevSettings& cfg, int buf,
                  uint32_t count) {
    if (!count) return;
    hipLaunchKernelGGL(k_shade, dim3(blocks(count, 256)), dim3(256), 0, s, sc, w, cfg, buf, count);
}
void launch_shadow(hipStream_t s, const DevScene& sc, const Wave& w, uint32_t count) {
    if (!count) return;
    hipLaunchKernelGGL(k_shadow, dim3(blocks(count, kTraceBlock)), dim3(kTraceBlock), 0, s, sc, w.s_o, w.s_d,
                       w.s_c, w.st_Le, w.st_La, count, w.stats);
}
void launch_accumulate(hipStream_t s, const Wave& w, float* image, uint32_t num_pixels, uint32_t first_iter,
                       uint32_t num_iters) {
    hipLaunchKernelGGL(k_accumulate, dim3(blocks(num_pixels, 256)), dim3(256), 0, s, w, image, num_pixels,
                       first_iter, num_iters);
}
void launch_trace_closest(hipStream_t s, const DevScene& sc, const float* rays, uint64_t n, float4* out) {
    if (!n) return;
    hipLaunchKernelGGL(k_trace_closest, dim3(blocks(n, kTraceBlock)), dim3(kTraceBlock), 0, s, sc, rays, n, out);
}
void launch_trace_occluded(hipStream_t s, const DevScene& sc, const float* rays, uint64_t n, uint8_t* out) {
    if (!n) return;
    hipLaunchKernelGGL(k_trace_occluded, dim3(blocks(n, kTraceBlock)), dim3(kTraceBlock), 0, s, sc, rays, n, out);
}
void launch_primary_rays(hipStream_t s, const DevScene& sc, uint32_t base_seed, uint32_t iteration, float* rays) {
    uint32_t np = (uint32_t)(sc.width * sc.height);
    hipLaunchKernelGGL(k_primary_rays, dim3(blocks(np, 256)), dim3(256), 0, s, sc, base_seed, iteration, rays);
}

}  // namespace sspt
