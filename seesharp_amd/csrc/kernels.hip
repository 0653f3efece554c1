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
#include <algorithm>

#include "adjoint.h"
#include "kernels.h"
#include "shading.h"

// Waves per SIMD the register allocator must leave room for (amdgpu_waves_per_eu):
// the traversal kernels are latency bound and gain from occupancy.
#ifndef SSPT_TRACE_WAVES
#define SSPT_TRACE_WAVES 6
#endif
#ifndef SSPT_SHADE_WAVES
#define SSPT_SHADE_WAVES 2
#endif
#ifndef SSPT_SHADE_BLOCK
#define SSPT_SHADE_BLOCK 256
#endif
#ifndef SSPT_SHADE_ADJ_WAVES
#define SSPT_SHADE_ADJ_WAVES 2
#endif

namespace sspt {

// threads per shade block: one queue atomic and one round of barriers per block and pass
constexpr int kShadeBlock = SSPT_SHADE_BLOCK;

// ------------------------------------------------------------ wave helpers -
// Adds the wave's count of `pred` to statistic `which`, on the stripe of this wave.
__device__ __forceinline__ uint32_t stat_stripe() {
    return (blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6)) & (uint32_t)(kStatStripes - 1);
}
__device__ __forceinline__ void wave_count(unsigned long long* stats, int which, bool pred) {
    uint64_t mask = __ballot(pred);
    if (mask == 0) return;
    int leader = __ffsll((unsigned long long)mask) - 1;
    if ((int)__lane_id() == leader)
        atomicAdd(&stats[stat_stripe() * kStatStride + which], (unsigned long long)__popcll(mask));
}
// Queue counters: extension queue b's length and the shadow rays appended
// with it, in one 64-bit word (low and high half).
__device__ __forceinline__ uint32_t queue_count(const Wave& w, int b) { return w.counters[2 * b]; }
__device__ __forceinline__ uint32_t shadow_count(const Wave& w, int b) { return w.counters[2 * b + 1]; }
__device__ __forceinline__ unsigned long long* queue_word(const Wave& w, int b) {
    return reinterpret_cast<unsigned long long*>(w.counters + 2 * b);
}

// Block-wide appends: one 64-bit atomic per block and call on a queue word,
// the block's `pa` entries (then its `pb` entries) to the half at `shift`
// (0: the extension queue, 32: the shadow rays), `pc` entries to the low half.
// Every thread of the block must make the same calls (block-uniform control
// flow); `lds` holds three uint32 per wave of the block plus two.
// HAS_B / HAS_C: whether the call has `pb` / `pc` entries at all.
template <bool HAS_B, bool HAS_C>
__device__ __forceinline__ uint3 block_append_word(unsigned long long* word, int shift, bool pa, bool pb, bool pc,
                                                   uint32_t* lds) {
    const uint64_t ma = __ballot(pa), mb = HAS_B ? __ballot(pb) : 0ull, mc = HAS_C ? __ballot(pc) : 0ull;
    const uint32_t wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
    const uint32_t lane = __lane_id();
    if (lane == 0) {
        lds[wave] = (uint32_t)__popcll(ma);
        if (HAS_B) lds[nw + wave] = (uint32_t)__popcll(mb);
        if (HAS_C) lds[2 * nw + wave] = (uint32_t)__popcll(mc);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t ab = 0;
        for (uint32_t k = 0; k < (HAS_B ? 2 : 1) * nw; ++k) {
            const uint32_t c = lds[k];
            lds[k] = ab;
            ab += c;
        }
        uint32_t cc = 0;
        if (HAS_C) {
            for (uint32_t k = 2 * nw; k < 3 * nw; ++k) {
                const uint32_t c = lds[k];
                lds[k] = cc;
                cc += c;
            }
        }
        const unsigned long long add = ((unsigned long long)ab << shift) + (unsigned long long)cc;
        const unsigned long long old = add ? atomicAdd(word, add) : 0ull;
        lds[3 * nw] = (uint32_t)(old >> shift);
        lds[3 * nw + 1] = (uint32_t)old;
    }
    __syncthreads();
    const uint64_t below = (1ull << lane) - 1ull;
    const uint3 r = make_uint3(lds[3 * nw] + lds[wave] + (uint32_t)__popcll(ma & below),
                               HAS_B ? lds[3 * nw] + lds[nw + wave] + (uint32_t)__popcll(mb & below) : 0u,
                               HAS_C ? lds[3 * nw + 1] + lds[2 * nw + wave] + (uint32_t)__popcll(mc & below) : 0u);
    __syncthreads();  // the next call reuses lds
    return r;
}

// ------------------------------------------------------------------ camera -
// PerspectiveCamera.GenerateRay (SeeSharp/Cameras/PerspectiveCamera.cs:64-102), lensRadius = 0:
// view = (2x/W - 1, 1 - 2y/H, 0); local = Transform(view, viewToCamera) (w = 1);
// dir = normalize(Transform((local, 0), cameraToWorld)); origin = camera position.
__device__ __forceinline__ void camera_ray(const DevCamera& c, float fx, float fy, f3* o, f3* d,
                                           f3* raw = nullptr) {
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
    *d = normalize_exact(dd);
    if (raw) *raw = dd;  // the unnormalised direction
}

// ------------------------------------------------------------ intersection -
struct RayPre {
    int kx, ky, kz;
    float sx, sy, sz;
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
    // Division-free prefilter: t = T / det lies in [tmin, tmax] only if the
    // sign-adjusted T lies in [tmin, tmax] * |det| widened by 2^-20 relative
    // (the rounded t below differs from T / det by < 2^-22 relative), so it
    // rejects only candidates the exact test rejects. A bound that underflows
    // could reject wrongly: the upper test needs a normal bound, and the lower
    // one only loosens when its bound underflows.
    const float ad = fabsf(det), Ts = det < 0.0f ? -T : T;
    const float lo = tmin * ad, hi = tmax * ad;
    if (Ts < lo - fabsf(lo) * 9.5367431640625e-07f) return false;
    if (hi >= 1e-37f && Ts > hi + hi * 9.5367431640625e-07f) return false;
    float rcp = 1.0f / det;
    float t = T * rcp;
    if (!(t >= tmin && t <= tmax)) return false;
    *t_out = t;
    *u_out = V * rcp;
    *v_out = W * rcp;
    return true;
}

// One traversal node's four child boxes (traversal node layout, scene_types.h):
// per axis the planes are origin + h * 2^e with fp16 offsets h, so the slab
// distance of a plane is h * (2^e * inv) + (origin * inv - q), one fma with an
// fp16 operand (v_fma_mix) per plane, where q is the ray's o * inv shifted by a
// margin per axis (BoxRay). The near plane of an axis is the lower one if the
// direction's component is positive, else the upper one (the sign bit decides,
// so -0 counts as negative). The margins make the computed near distance no
// larger and the far distance no smaller than the exact slab distances of the
// decoded planes, which enclose the children's boxes, so a box the ray meets
// within [tmin, tmax] is never culled. Returns the hit slots as a bit mask and
// their entry distances in tc (each in [tmin, tmax] for a hit slot).
__device__ __forceinline__ float h16(uint32_t w, int upper) {
    return (float)__builtin_bit_cast(_Float16, (uint16_t)(upper ? w >> 16 : w & 0xffffu));
}
template <class BR>
__device__ __forceinline__ uint32_t child_boxes(const uint32_t (&w)[kTravNodeWords], const BR& box, float tmin,
                                                float tmax, float (&tc)[4]) {
    const uint32_t meta = w[3];
    const float ox = __uint_as_float(w[0]), oy = __uint_as_float(w[1]), oz = __uint_as_float(w[2]);
    const float sx = ldexpf(box.inv.x, (int)(int8_t)(meta & 0xffu));
    const float sy = ldexpf(box.inv.y, (int)(int8_t)((meta >> 8) & 0xffu));
    const float sz = ldexpf(box.inv.z, (int)(int8_t)((meta >> 16) & 0xffu));
    const float bnx = fmaf(ox, box.inv.x, -box.qn.x), bfx = fmaf(ox, box.inv.x, -box.qf.x);
    const float bny = fmaf(oy, box.inv.y, -box.qn.y), bfy = fmaf(oy, box.inv.y, -box.qf.y);
    const float bnz = fmaf(oz, box.inv.z, -box.qn.z), bfz = fmaf(oz, box.inv.z, -box.qf.z);
    const bool gx = signbit(box.inv.x), gy = signbit(box.inv.y), gz = signbit(box.inv.z);
    const uint32_t nx0 = gx ? w[6] : w[4], nx1 = gx ? w[7] : w[5], fx0 = gx ? w[4] : w[6], fx1 = gx ? w[5] : w[7];
    const uint32_t ny0 = gy ? w[10] : w[8], ny1 = gy ? w[11] : w[9], fy0 = gy ? w[8] : w[10], fy1 = gy ? w[9] : w[11];
    const uint32_t nz0 = gz ? w[14] : w[12], nz1 = gz ? w[15] : w[13], fz0 = gz ? w[12] : w[14],
                   fz1 = gz ? w[13] : w[15];
    uint32_t hit = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const uint32_t nxw = k < 2 ? nx0 : nx1, fxw = k < 2 ? fx0 : fx1;
        const uint32_t nyw = k < 2 ? ny0 : ny1, fyw = k < 2 ? fy0 : fy1;
        const uint32_t nzw = k < 2 ? nz0 : nz1, fzw = k < 2 ? fz0 : fz1;
        const int u = k & 1;
        const float tn = fmaxf(fmaxf(fmaf(h16(nxw, u), sx, bnx), fmaf(h16(nyw, u), sy, bny)),
                               fmaxf(fmaf(h16(nzw, u), sz, bnz), tmin));
        const float tf = fminf(fminf(fmaf(h16(fxw, u), sx, bfx), fmaf(h16(fyw, u), sy, bfy)),
                               fminf(fmaf(h16(fzw, u), sz, bfz), tmax));
        tc[k] = tn;
        hit |= (uint32_t)(tn <= tf) << k;
    }
    return hit;
}

// Per-ray constants of child_boxes. Let M = max(|o|, R), R the scene's largest
// vertex coordinate magnitude: node origins lie within R and decoded planes P
// within 1.004 R (an upper plane rounds up by at most 2^-9 of its node's
// extent, which is below 2 R). With A = 2^-20 * M * |inv| per axis,
// qn = o * inv + A and qf = o * inv - A (each one rounding, u = 2^-24). A
// computed near distance h * s + (origin * inv - qn), s = 2^e * inv exact,
// differs from (P - o) * inv - A by at most 5.1 u M |inv| (the roundings of qn,
// of the base and of the fma), and (P - o) * inv from the exact (P - o) / d by
// at most 2.1 u M |inv| (the rounded reciprocal), so it is at least 8 u M |inv|
// below the exact near distance, and a far distance as far above the exact one.
// Direction components below 2^-40 of the largest one are replaced by +-2^-40
// of it (sign kept): the altered ray stays within t * 2^-40 * |d|max of the
// true one along that axis, and where the true ray can be inside a box
// (t * |d|max <= 2 M) that is far below the margin's 2^-20 M, so the
// substitution keeps the test conservative. A direction shorter than 2^-60 M
// (or not finite) gets inv = 0, qn = +inf, qf = -inf: every box is entered.
// Otherwise M |inv| <= 2^100 and every distance is finite.
struct BoxRay {
    f3 inv, qn, qf;
};
__device__ __forceinline__ BoxRay box_ray(float abs_max, f3 o, f3 d) {
    const float dmax = fmaxf(fmaxf(fabsf(d.x), fabsf(d.y)), fabsf(d.z));
    const float lo = dmax * 0x1p-40f;
    const float cx = fabsf(d.x) < lo ? copysignf(lo, d.x) : d.x;
    const float cy = fabsf(d.y) < lo ? copysignf(lo, d.y) : d.y;
    const float cz = fabsf(d.z) < lo ? copysignf(lo, d.z) : d.z;
    BoxRay b;
    const float mm = fmaxf(abs_max, fmaxf(fmaxf(fabsf(o.x), fabsf(o.y)), fabsf(o.z)));
    if (!(dmax > mm * 0x1p-60f)) {
        b.inv = mk3(0.0f, 0.0f, 0.0f);
        b.qn = mk3(INFINITY, INFINITY, INFINITY);
        b.qf = mk3(-INFINITY, -INFINITY, -INFINITY);
        return b;
    }
    b.inv = mk3(1.0f / cx, 1.0f / cy, 1.0f / cz);
    const float m = mm * 0x1p-20f;
    const f3 a = mk3(m * fabsf(b.inv.x), m * fabsf(b.inv.y), m * fabsf(b.inv.z));
    b.qn = mk3(fmaf(o.x, b.inv.x, a.x), fmaf(o.y, b.inv.y, a.y), fmaf(o.z, b.inv.z, a.z));
    b.qf = mk3(fmaf(o.x, b.inv.x, -a.x), fmaf(o.y, b.inv.y, -a.y), fmaf(o.z, b.inv.z, -a.z));
    return b;
}

// The watertight t carries an absolute error of a few ulps of the coordinate
// magnitudes (divided by |d|), not of t itself. Box culling against the best
// hit so far uses [tmin - slack, tmax + slack] with slack = 2^-16 * max(|o|,
// scene extent) / |d|, so a box is never culled by rounding against a triangle
// distance that is equal in exact arithmetic.
__device__ __forceinline__ float ray_slack(float abs_max, f3 o, f3 d) {
    float m = fmaxf(abs_max, fmaxf(fmaxf(fabsf(o.x), fabsf(o.y)), fabsf(o.z)));
    return m * 1.52587890625e-05f / length(d);
}

struct HitOut {
    float t, u, v;
    int li;   // leaf-order triangle index, -1 = miss
    int gid;  // global triangle id
};

// Per-lane traversal stack of (node, entry distance): the first kStackLds
// entries in LDS ([depth][lane] layout, conflict free), the rest in a private
// array that only deep trees reach. The private arrays are separate locals of
// the caller so that the traversal state around them stays in registers.
constexpr int kStackSpill = kStackTotal - kStackLds;
// KEEP_T: entries carry the node's entry distance (closest-hit traversal culls
// popped nodes behind the best hit); an any-hit traversal keeps nodes only.
template <int TB, bool KEEP_T>  // TB: threads per block, the stride of a lane's LDS entries
struct TraceStack {
    int* lds_node;    // + threadIdx.x, stride TB
    float* lds_t;
    int* spill_node;  // private arrays of kStackSpill entries, owned by the caller
    float* spill_t;
    int sp;
    __device__ __forceinline__ void push(int node, float t) {
        // The builder rejects trees needing more than kStackTotal entries; the
        // guard keeps memory safe regardless.
        if (sp >= kStackTotal) return;
        if (sp < kStackLds) {
            lds_node[sp * TB] = node;
            if (KEEP_T) lds_t[sp * TB] = t;
        } else {
            spill_node[sp - kStackLds] = node;
            if (KEEP_T) spill_t[sp - kStackLds] = t;
        }
        ++sp;
    }
    // push when the lane's LDS part has room (the caller checked sp < kStackLds)
    __device__ __forceinline__ void push_lds(int node, float t) {
        lds_node[sp * TB] = node;
        if (KEEP_T) lds_t[sp * TB] = t;
        ++sp;
    }
    __device__ __forceinline__ void pop(int* node, float* t) {
        --sp;
        if (sp < kStackLds) {
            *node = lds_node[sp * TB];
            *t = KEEP_T ? lds_t[sp * TB] : 0.0f;
        } else {
            *node = spill_node[sp - kStackLds];
            *t = KEEP_T ? spill_t[sp - kStackLds] : 0.0f;
        }
    }
};

__device__ __forceinline__ void cswap(float& ta, int& na, float& tb, int& nb) {
    bool sw = tb < ta;
    float t = sw ? tb : ta;
    int n = sw ? nb : na;
    tb = sw ? ta : tb;
    nb = sw ? na : nb;
    ta = t;
    na = n;
}

// 4-wide BVH traversal of one ray. Closest hit visits hit children nearest
// first and skips popped nodes whose entry distance is beyond the best hit;
// ANY = true stops at the first hit (shadow rays).
//
// The triangles of hit leaves are not tested in the node step that finds
// them: the leaf goes to a per-lane FIFO of kPendCap entries and the lane
// keeps traversing. `traverse` runs triangle steps (one queued leaf per lane)
// when most of the wave's live lanes have a leaf queued, when a lane's FIFO is
// full, or when no lane has nodes left. A wave then spends its triangle tests
// on many lanes at once instead of one node step's few leaves (the SIMD cost
// of a traversal is its node steps plus its triangle steps, each paid by the
// whole wave). The result is the same: a leaf is queued only while its box
// meets [tmin, best], the closest hit is decided by distance and then by the
// lower global id whatever the test order, and an any-hit query only asks
// whether some triangle is hit.
template <bool ANY, int TB>
struct Tracer {
    f3 o;
    RayPre p;
    BoxRay box;
    float tmin, slack, tmin_box;
    HitOut best;
    TraceStack<TB, !ANY> st;
    int* pend;     // FIFO of queued leaves in LDS (+ threadIdx.x, stride TB)
    uint32_t ph;    // FIFO head
    uint32_t pn;    // FIFO length
    int node;       // next node to visit, -1 when none is left
    bool done;      // any-hit found (or an invalid node index)

    __device__ __forceinline__ void start(int root, float abs_max, f3 o_, f3 d, float tmin_, float tmax) {
        o = o_;
        p = precompute(d);
        box = box_ray(abs_max, o_, d);
        slack = ray_slack(abs_max, o_, d);
        tmin = tmin_;
        tmin_box = tmin_ - slack;
        best = HitOut{tmax, 0.0f, 0.0f, -1, 0x7fffffff};
        st.sp = 0;
        ph = 0;
        pn = 0;
        node = root;
        done = false;
    }

    // Lanes still working: nodes to visit or leaves queued.
    __device__ __forceinline__ bool live() const { return !done && (node >= 0 || pn > 0); }
    // A node step may queue four leaves.
    __device__ __forceinline__ bool can_step() const { return !done && node >= 0 && pn <= (uint32_t)(kPendCap - 4); }

    // Triangles tri .. tri + n - 1 of one leaf.
    __device__ __forceinline__ void test_leaf(const DevScene& S, int tri, int n) {
        for (int j = 0; j < n; ++j, ++tri) {
            float4 a = S.tri_pos[3 * tri], b = S.tri_pos[3 * tri + 1], c = S.tri_pos[3 * tri + 2];
            float t, u, v;
            if (intersect_tri(o, p, tmin, best.t, f3_of(a), f3_of(b), f3_of(c), &t, &u, &v)) {
                int gid = __float_as_int(a.w);
                if (ANY || t < best.t || gid < best.gid) {
                    best.t = t;
                    best.u = u;
                    best.v = v;
                    best.li = tri;
                    best.gid = gid;
                    if (ANY) {
                        done = true;
                        return;
                    }
                }
            }
        }
    }

    // Box tests of the current node's children: hit leaves are queued (a leaf
    // of more than four triangles, which only the build's depth bound makes,
    // is tested here), hit inner children are pushed nearest first and the
    // nearest becomes the next node, else the stack is popped. The node comes
    // through the scalar data cache when every stepping lane of the wave is at
    // it, else with five 16-byte vector loads per lane: the traversal's time
    // follows its vector memory instructions (a 64-lane scattered load costs
    // the texture units about the same whatever its width), not the bytes.
    __device__ __forceinline__ void node_step(const DevScene& S) {
        if ((unsigned)node >= (unsigned)S.num_nodes) {  // never taken for a valid tree
            done = true;
            return;
        }
        // finite, so that a hit child's entry distance (<= tmax_box) is finite
        const float tmax_box = fminf(best.t + slack, 3.402823466e38f);
        float tc[4];
        uint32_t hit, kinds;
        int cw[4];
        const int nu = __builtin_amdgcn_readfirstlane(node);
        if (__ballot(node == nu) == __ballot(true)) {
            // every stepping lane of the wave visits the same node: scalar loads
            // (the scalar data cache), no vector memory instructions
            typedef const __attribute__((address_space(4))) uint32_t ConstU32;
            ConstU32* q = (ConstU32*)(S.qnodes + (size_t)nu * kTravNodeWords);
            uint32_t w[kTravNodeWords];
#pragma unroll
            for (int i = 0; i < kTravNodeWords; ++i) w[i] = q[i];
            hit = child_boxes(w, box, tmin_box, tmax_box, tc);
            kinds = w[3] >> 24;
            cw[0] = (int)w[16];
            cw[1] = (int)w[17];
            cw[2] = (int)w[18];
            cw[3] = (int)w[19];
        } else {
            // 80 B per node; scene creation bounds node * 80 below 2^32
            const char* B = reinterpret_cast<const char*>(S.qnodes);
            const uint32_t nb = (uint32_t)node * 80u;
            uint32_t w[kTravNodeWords];
#pragma unroll
            for (int i = 0; i < 5; ++i) {
                const uint4 v = *reinterpret_cast<const uint4*>(B + nb + 16u * i);
                w[4 * i] = v.x;
                w[4 * i + 1] = v.y;
                w[4 * i + 2] = v.z;
                w[4 * i + 3] = v.w;
            }
            hit = child_boxes(w, box, tmin_box, tmax_box, tc);
            kinds = w[3] >> 24;
            cw[0] = (int)w[16];
            cw[1] = (int)w[17];
            cw[2] = (int)w[18];
            cw[3] = (int)w[19];
        }
        const uint32_t leafk = kinds & 15u, innerk = kinds >> 4;
        uint32_t leaves = hit & leafk & ~innerk;  // hit leaves of at most four triangles
        uint32_t bigs = hit & leafk & innerk;     // hit leaves of more triangles
        const uint32_t inner_hit = hit & innerk & ~leafk;
        bool hc[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) hc[k] = (inner_hit >> k) & 1u;
        while (leaves) {
            const int k = __builtin_ctz(leaves);
            leaves &= leaves - 1u;
            // the child word is the FIFO entry: first triangle (< 2^30) | (count - 1) << 30
            pend[((ph + pn) & (uint32_t)(kPendCap - 1)) * TB] = k == 0 ? cw[0] : k == 1 ? cw[1] : k == 2 ? cw[2] : cw[3];
            ++pn;
        }
        if (__ballot(bigs != 0)) {  // only trees with leaves of more than four triangles
            while (bigs) {
                const int k = __builtin_ctz(bigs);
                bigs &= bigs - 1u;
                const int b = k == 0 ? cw[0] : k == 1 ? cw[1] : k == 2 ? cw[2] : cw[3];
                test_leaf(S, S.big_leaves[2 * b], S.big_leaves[2 * b + 1]);
                if (ANY && done) return;
            }
        }
        // inner children, nearest first. A hit child's entry distance lies in
        // [tmin_box, tmax_box] (finite), so every hit child sorts before every
        // other slot (key +inf).
        float t0 = hc[0] ? tc[0] : INFINITY;
        float t1 = hc[1] ? tc[1] : INFINITY;
        float t2 = hc[2] ? tc[2] : INFINITY;
        float t3 = hc[3] ? tc[3] : INFINITY;
        int n0 = cw[0], n1 = cw[1], n2 = cw[2], n3 = cw[3];
        int nhit = __builtin_popcount(inner_hit);
        if (nhit > 0) {
            if (!ANY && __ballot(nhit > 1) == 0) {
                // no lane of the wave has two hit children: visit the one, no sort
                const uint32_t k0 = hc[0] ? (uint32_t)n0 : 0xffffffffu, k1 = hc[1] ? (uint32_t)n1 : 0xffffffffu;
                const uint32_t k2 = hc[2] ? (uint32_t)n2 : 0xffffffffu, k3 = hc[3] ? (uint32_t)n3 : 0xffffffffu;
                node = (int)min(min(k0, k1), min(k2, k3));
                return;
            }
            if (!ANY) {
                cswap(t0, n0, t1, n1);
                cswap(t2, n2, t3, n3);
                cswap(t0, n0, t2, n2);
                cswap(t1, n1, t3, n3);
                cswap(t1, n1, t2, n2);
            } else {
                // any-hit: an occluder found in any order ends the ray, so the hit
                // children only move to the front (node indices < 2^25 sort
                // before the absent key), in index order
                uint32_t k0 = hc[0] ? (uint32_t)n0 : 0xffffffffu, k1 = hc[1] ? (uint32_t)n1 : 0xffffffffu;
                uint32_t k2 = hc[2] ? (uint32_t)n2 : 0xffffffffu, k3 = hc[3] ? (uint32_t)n3 : 0xffffffffu;
                uint32_t a = min(k0, k1), b = max(k0, k1), c = min(k2, k3), d = max(k2, k3);
                k0 = min(a, c);
                k2 = max(a, c);
                k1 = min(b, d);
                k3 = max(b, d);
                n0 = (int)k0;
                n1 = (int)min(k1, k2);
                n2 = (int)max(k1, k2);
                n3 = (int)k3;
            }
            if (__ballot(st.sp + 3 > kStackLds) == 0) {  // every lane's pushes fit in LDS
                if (nhit > 3) st.push_lds(n3, t3);
                if (nhit > 2) st.push_lds(n2, t2);
                if (nhit > 1) st.push_lds(n1, t1);
            } else {
                if (nhit > 3) st.push(n3, t3);
                if (nhit > 2) st.push(n2, t2);
                if (nhit > 1) st.push(n1, t1);
            }
            node = n0;
            return;
        }
        // pop the next node still in front of the best hit
        while (st.sp > 0) {
            float tn;
            st.pop(&node, &tn);
            if (ANY || tn <= best.t + slack) return;
        }
        node = -1;
    }

    // Tests the oldest queued leaf.
    __device__ __forceinline__ void tri_step(const DevScene& S) {
        const uint32_t e = (uint32_t)pend[(ph & (uint32_t)(kPendCap - 1)) * TB];
        ++ph;
        --pn;
        test_leaf(S, (int)(e & 0x3fffffffu), (int)(e >> 30) + 1);
    }
};

// Runs one ray per lane to its end. The active lanes of a wave call it
// together: the choice between node and triangle steps is made per wave.
template <bool ANY, int TB = kTraceBlock>
__device__ __forceinline__ HitOut traverse(const DevScene& S, int root, float abs_max, f3 o, f3 d, float tmin,
                                           float tmax, int* lds_node, float* lds_t) {
    int spill_node[kStackSpill];
    float spill_t[kStackSpill];
    Tracer<ANY, TB> tr;
    tr.st.lds_node = lds_node;
    tr.st.lds_t = lds_t;
    tr.st.spill_node = spill_node;
    tr.st.spill_t = spill_t;
    tr.pend = lds_node + kStackLds * TB;
    tr.start(root, abs_max, o, d, tmin, tmax);
    for (;;) {
        if (tr.can_step()) tr.node_step(S);
        const bool live = tr.live();
        const uint64_t act = __ballot(live);
        if (act == 0) break;
        const uint64_t queued = __ballot(live && tr.pn > 0);
        const uint64_t walking = __ballot(live && tr.node >= 0);
        const uint64_t can = __ballot(tr.can_step());
        // triangle step: kPendPercent % of the live lanes have a leaf queued,
        // a lane cannot step for a full FIFO, or no lane has a node left
        if (queued != 0 &&
            (can == 0 || walking != can ||
             100 * __popcll(queued) >= (ANY ? kPendPercentAny : kPendPercent) * __popcll(act))) {
            if (live && tr.pn > 0) tr.tri_step(S);
        }
    }
    return tr.best;
}

// ---------------------------------------------------------------- kernels --
// Every wave kernel is persistent: a fixed grid loops over the device-side
// length of its queue, so a whole wave runs without host synchronisation and
// can be captured in a graph.

__device__ __forceinline__ const WaveDesc& cur_wave(const Wave& w) { return w.waves[w.counters[kCntWave]]; }

__global__ void k_wave_begin(Wave w) {
    const WaveDesc& d = cur_wave(w);
    w.counters[kCntQueue0] = d.count;
    w.counters[kCntShadow0] = 0;
    w.counters[kCntQueue1] = 0;
    w.counters[kCntShadow1] = 0;
}

__global__ void k_wave_end(Wave w) { w.counters[kCntWave] += 1; }

__global__ void k_bounce_reset(Wave w, int next) {
    w.counters[2 * next] = 0;
    w.counters[2 * next + 1] = 0;
}

// Scene of world-wide path g: last scene whose path offset is <= g.
__device__ __forceinline__ int scene_of(const DevScene& S, uint64_t g) {
    int lo = 0, hi = S.num_scenes - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (S.path_offsets[mid] <= g) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// Pixel of the k-th path of an iteration. The rows are walked in bands of
// kTileRows rows and each band column by column, so the 64 consecutive paths
// of a wavefront cover an 8 x 8 tile of the image: their camera rays, and the
// rays of the first bounces that keep their queue order, share more BVH nodes
// than a 64 x 1 strip of one row does. A bijection onto the W x H pixels for
// any W, H (the last band is H mod 8 rows high). The order only decides which
// slot a pixel's path occupies: every path keeps its own RNG stream.
#ifndef SSPT_TILE_ROWS
#define SSPT_TILE_ROWS 8
#endif
constexpr uint32_t kTileRows = SSPT_TILE_ROWS;
__device__ __forceinline__ uint32_t path_pixel(uint32_t k, uint32_t W, uint32_t H) {
    if (kTileRows <= 1) return k;
    const uint32_t band = k / (kTileRows * W);
    const uint32_t r = k - band * kTileRows * W;
    const uint32_t bh = min(kTileRows, H - band * kTileRows);
    const uint32_t col = r / bh;
    return (band * kTileRows + (r - col * bh)) * W + col;
}

// RNG(BaseSeed, row * W + col, iteration) (PathTracer.cs:236-237), pixel jitter
// (:289-290) and camera ray (:291) of world-wide path g.
struct PathStart {
    int scene;
    uint32_t pix;
    Rng rng;
    f3 o, dir;
};
__device__ __forceinline__ PathStart start_path(const DevScene& S, uint64_t g) {
    PathStart p;
    p.scene = scene_of(S, g);
    const SceneRec& R = S.recs[p.scene];
    uint64_t k = g - R.path_off;
    uint32_t it = (uint32_t)(k / R.num_pixels);
    p.pix = path_pixel((uint32_t)(k - (uint64_t)it * R.num_pixels), (uint32_t)R.cam.width, (uint32_t)R.cam.height);
    uint32_t row = p.pix / (uint32_t)R.cam.width;
    uint32_t col = p.pix - row * (uint32_t)R.cam.width;
    p.rng = Rng{hash_seed(R.base_seed, p.pix, R.first_iter + it)};
    f2 off = p.rng.next_float2();
    camera_ray(R.cam, (float)col + off.x, (float)row + off.y, &p.o, &p.dir);
    return p;
}

// The initial PathState (PathTracer.cs:293-302) of the path in `slot`, and
// its entry in extension queue 0 (all but the ray origin).
__device__ __forceinline__ void init_path(const DevScene& S, const Wave& w, uint32_t slot, const PathStart& p) {
    w.q_d[0][slot] = f4_of(p.dir, INFINITY);
    w.q_slot[0][slot] = slot | ((uint32_t)p.scene << kSlotBits);
    w.st_pix[slot] = p.pix | ((uint32_t)p.scene << kSlotBits);
    w.q_rng[0][slot] = p.rng.state;
    w.q_thr[0][slot] = make_float4(1.0f, 1.0f, 1.0f, 0.0f);
    w.q_prev[0][slot] = make_float4(0.0f, 0.0f, 0.0f, __int_as_float(1));
    w.q_L[0][slot] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    w.st_Le[slot] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    w.st_La[slot] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if (S.recs[p.scene].enable_denoiser) {
        w.st_alb[slot] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        w.st_nrm[slot] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    }
}

// The first extension ray and the initial PathState of every path of the wave.
__global__ void __launch_bounds__(256) k_generate(DevScene S, Wave w) {
    const WaveDesc d = cur_wave(w);
    for (uint32_t slot = blockIdx.x * blockDim.x + threadIdx.x; slot < d.count; slot += gridDim.x * blockDim.x) {
        PathStart p = start_path(S, d.start + slot);
        w.q_o[0][slot] = f4_of(p.o, 0.0f);
        init_path(S, w, slot, p);
    }
}

// Closest hits of the extension queue `buf`, over blocks [block, ...) of a
// grid of `nblocks` blocks that stride over the device-side queue length.
template <int TB>
__device__ __forceinline__ void extend_rays(const DevScene& S, const Wave& w, int buf, uint32_t block,
                                            uint32_t nblocks, int* stack_node, float* stack_t) {
    const uint32_t count = queue_count(w, buf);
    const float4* __restrict__ q_o = w.q_o[buf];
    const float4* __restrict__ q_d = w.q_d[buf];
    const uint32_t* __restrict__ q_slot = w.q_slot[buf];
    for (uint32_t base = block * blockDim.x; base < count; base += nblocks * blockDim.x) {
        uint32_t i = base + threadIdx.x;
        bool active = i < count;
        HitOut h{0.0f, 0.0f, 0.0f, -1, 0};
        if (active) {
            float4 o4 = q_o[i], d4 = q_d[i];
            const SceneRec& R = S.recs[q_slot[i] >> kSlotBits];
            h = traverse<false, TB>(S, R.root, R.abs_max, f3_of(o4), f3_of(d4), o4.w, d4.w, stack_node + threadIdx.x,
                                    stack_t + threadIdx.x);
            w.hits[i] = h.li >= 0 ? make_float4(h.t, h.u, h.v, __int_as_float(h.li))
                                  : make_float4(INFINITY, 0.0f, 0.0f, __int_as_float(-1));
        }
        wave_count(w.stats, kStatRays, active);
        wave_count(w.stats, kStatHits, active && h.li >= 0);
    }
}

// Any-hit tests of the shadow queue; a visible ray adds its contribution to
// its path's next-event sum (and, with ADJ, its own gradient term).
template <bool ADJ, int TB>
__device__ __forceinline__ void shadow_rays(const DevScene& S, const Wave& w, int slot, uint32_t block,
                                            uint32_t nblocks,
                                            int* stack_node, float* stack_t) {
    const uint32_t count = shadow_count(w, slot);
    for (uint32_t base = block * blockDim.x; base < count; base += nblocks * blockDim.x) {
        uint32_t i = base + threadIdx.x;
        bool active = i < count;
        bool occluded = false;
        if (active) {
            float4 o4 = w.s_o[i], d4 = w.s_d[i];
            float4 c = w.s_c[i];
            uint32_t tag = (uint32_t)__float_as_int(c.w);
            const SceneRec& R = S.recs[tag >> (kSlotBits + 1)];
            HitOut h = traverse<true, TB>(S, R.root, R.abs_max, f3_of(o4), f3_of(d4), o4.w, d4.w,
                                          stack_node + threadIdx.x, stack_t + threadIdx.x);
            occluded = h.li >= 0;
            if (!occluded) {
                // the contribution is non-zero by construction
                float4* dst = (tag & 1u) ? w.st_La : w.st_Le;
                uint32_t slot = (tag >> 1) & kSlotMask;
                atomicAdd(&dst[slot].x, c.x);
                atomicAdd(&dst[slot].y, c.y);
                atomicAdd(&dst[slot].z, c.z);
                if (ADJ) {
                    // the vertex's own term: d(contribution)/d(parameters), final once visible
                    int2 t = w.s_t[i];
                    float4 g = w.s_g[i];
                    if (t.x >= 0) {
                        atomicAdd(&S.grad[t.x], g.x);
                        atomicAdd(&S.grad[t.x + 1], g.y);
                        atomicAdd(&S.grad[t.x + 2], g.z);
                    }
                    if (t.y >= 0) atomicAdd(&S.grad[t.y], g.w);
                }
            }
        }
        wave_count(w.stats, kStatShadowTraced, active);
        wave_count(w.stats, kStatOccluded, active && occluded);
    }
}

// k_generate and the closest hits of its queue in one pass: each thread starts
// its path (RNG, jitter, camera ray), writes the path state and traces the
// camera ray at once, so the ray never goes through the queue.
template <int TB>
__global__ void __launch_bounds__(TB) __attribute__((amdgpu_waves_per_eu(SSPT_TRACE_WAVES)))
k_extend_primary(DevScene S, Wave w) {
    __shared__ int stack_node[(kStackLds + kPendCap) * TB];
    __shared__ float stack_t[kStackLds * TB];
    const WaveDesc d = cur_wave(w);
    for (uint32_t base = blockIdx.x * blockDim.x; base < d.count; base += gridDim.x * blockDim.x) {
        const uint32_t slot = base + threadIdx.x;
        const bool active = slot < d.count;
        HitOut h{0.0f, 0.0f, 0.0f, -1, 0};
        if (active) {
            const PathStart p = start_path(S, d.start + slot);
            init_path(S, w, slot, p);
            const SceneRec& R = S.recs[p.scene];
            h = traverse<false, TB>(S, R.root, R.abs_max, p.o, p.dir, 0.0f, INFINITY, stack_node + threadIdx.x,
                                stack_t + threadIdx.x);
            w.hits[slot] = h.li >= 0 ? make_float4(h.t, h.u, h.v, __int_as_float(h.li))
                                     : make_float4(INFINITY, 0.0f, 0.0f, __int_as_float(-1));
        }
        wave_count(w.stats, kStatRays, active);
        wave_count(w.stats, kStatHits, active && h.li >= 0);
    }
}

template <int TB>
__global__ void __launch_bounds__(TB) __attribute__((amdgpu_waves_per_eu(SSPT_TRACE_WAVES)))
k_extend(DevScene S, Wave w, int buf) {
    __shared__ int stack_node[(kStackLds + kPendCap) * TB];
    __shared__ float stack_t[kStackLds * TB];
    extend_rays<TB>(S, w, buf, blockIdx.x, gridDim.x, stack_node, stack_t);
}

template <bool ADJ, int TB>
__global__ void __launch_bounds__(TB) __attribute__((amdgpu_waves_per_eu(SSPT_TRACE_WAVES)))
k_shadow(DevScene S, Wave w, int slot) {
    __shared__ int stack_node[(kStackLds + kPendCap) * TB];
    __shared__ float stack_t[kStackLds * TB];
    shadow_rays<ADJ, TB>(S, w, slot, blockIdx.x, gridDim.x, stack_node, stack_t);
}

// One launch for the two traversals that are ready after a bounce's shade:
// the continuation rays of the next bounce (blocks [0, extend_blocks)) and
// the bounce's shadow rays (the remaining blocks). Neither reads what the
// other writes, and a single grid lets the tail of one overlap the other.
template <bool ADJ, int TB>
__global__ void __launch_bounds__(TB) __attribute__((amdgpu_waves_per_eu(SSPT_TRACE_WAVES)))
k_trace(DevScene S, Wave w, int buf, uint32_t extend_blocks) {
    __shared__ int stack_node[(kStackLds + kPendCap) * TB];
    __shared__ float stack_t[kStackLds * TB];
    if (blockIdx.x < extend_blocks)
        extend_rays<TB>(S, w, buf, blockIdx.x, extend_blocks, stack_node, stack_t);
    else
        shadow_rays<ADJ, TB>(S, w, buf, blockIdx.x - extend_blocks, gridDim.x - extend_blocks, stack_node, stack_t);
}

// TinyEmbree SpawnRay / IsOccluded offset rule: move along the face normal by
// the error offset, to the side the ray leaves on.
__device__ __forceinline__ f3 offset_point(const Surf& s, f3 dir) {
    float sign = dot(dir, s.ng) < 0.0f ? -1.0f : 1.0f;
    return s.pos + sign * s.err * s.ng;
}

__device__ __forceinline__ uint32_t shadow_tag(uint32_t slot, uint32_t scene, uint32_t kind) {
    return (scene << (kSlotBits + 1)) | (slot << 1) | kind;
}

// Texel float offsets of the shaded vertex's textured parameters (-1 = constant).
__device__ __forceinline__ int2 vertex_texels(const DevScene& S, const Shader& sh, f2 uv) {
    int2 t = make_int2(-1, -1);
    if (sh.m.base_tex >= 0) {
        DevTexture tx = S.textures[sh.m.base_tex];
        t.x = (int)(tx.offset + 3ull * (uint64_t)texel_index(tx, uv));
    }
    if (sh.m.kind == 1 && sh.m.rough_tex >= 0) {
        DevTexture tx = S.textures[sh.m.rough_tex];
        t.y = (int)(tx.offset + (uint64_t)texel_index(tx, uv));
    }
    return t;
}

// Emitter.EmittedRadiance at a point with shading normal ns, towards dir:
// DiffuseEmitter (DiffuseEmitter.cs:33-37) is one-sided on ns; GlossyEmitter
// (GlossyEmitter.cs:36-40) scales by cos^exponent * (exponent + 1) / (2 pi).
__device__ __forceinline__ f3 emitted(const DevEmitter& e, f3 ns, f3 dir) {
    if (e.kind == kEmitterGlossy) {
        float cosine = sdiv(dot(ns, dir), length(dir));
        if (cosine < 0.0f) return gray(0.0f);
        return mk3(e.r, e.g, e.b) * powf(cosine, e.exponent) * e.norm;
    }
    return dot(ns, dir) < 0.0f ? gray(0.0f) : mk3(e.r, e.g, e.b);
}

// What ReplayPixel (PathTracer.cs:61-69) records of one vertex: the values the
// reference's PathGraph nodes carry (PathGraph.cs BSDFSampleNode,
// NextEventNode, BackgroundNode). Filled only by the replay kernel; every other
// user leaves PathVertex::log null and the recording compiles away.
struct ReplayLog {
    bool hit_light = false, hit_background = false;  // emitter hit / background hit at depth >= MinDepth
    f3 hit_emission;
    float hit_mis = 0.0f;
    bool env_ok = false, env_nonzero = false;  // pdfs non-zero / contribution non-black
    f3 env_dir, env_emission, env_bc;
    float env_pdf = 0.0f, env_mis = 0.0f;
    float4 env_o, env_d;                       // the environment shadow ray
    bool area_ok = false, area_nonzero = false;
    f3 area_pos, area_emission, area_bc;
    float area_pdf = 0.0f, area_mis = 0.0f;
    float4 area_o, area_d;                     // the area-light shadow ray
    f3 scatter;                                // BSDF sample weight / survival probability
};

// One vertex of EstimateIncidentRadiance (PathTracer.cs:315-382) for one path:
// the hit of its extension ray (background / emitter with MIS, Russian
// roulette), its next-event rays (environment, then area light, per shadow
// ray) and its continuation. The wavefront shade kernel queues the rays a
// vertex emits; the single-kernel tracer traces them at once. With ADJ the
// vertex also computes the adjoint's per-ray gradients.
// Emitters a shade block stages in LDS (single-scene worlds with at most this many).
constexpr int kLdsEmitters = 64;
constexpr int kEmitterWords = 9;  // float4: DevEmitter (2), tri_pos (3), tri_aux (4)
static_assert(sizeof(DevEmitter) == 32, "DevEmitter is two float4");

template <bool ADJ>
struct PathVertex {
    const DevScene S;
    const SceneRec& R;
    Env E;
    uint32_t slot, sc;
    // path state (PathState, PathTracer.cs:143-194)
    Rng rng{0};
    f3 prefix = {0.0f, 0.0f, 0.0f};
    f3 L = {0.0f, 0.0f, 0.0f};
    f3 dir = {0.0f, 0.0f, 0.0f};  // the extension ray's direction
    f3 prev_pos = {0.0f, 0.0f, 0.0f};
    float prev_pdf = 0.0f;
    int depth = 0;
    // this vertex
    Surf s;
    Shader sh;
    float q = 1.0f;
    bool cont = false, do_nee = false;
    bool surface_hit = false;       // the extension ray hit a surface (s, sh are valid)
    int2 vtex = {-1, -1};           // ADJ: texel offsets of the vertex's textured parameters
    // the scene's emitters staged in LDS (kEmitterWords float4 each: the record,
    // then its triangle's 3 tri_pos and 4 tri_aux entries), or null
    const float4* em_tab = nullptr;
    f3 next_dir = {0.0f, 0.0f, 0.0f};
    ReplayLog* log = nullptr;  // ReplayPixel only

    __device__ PathVertex(const DevScene& S_, const SceneRec& R_, uint32_t slot_, uint32_t sc_)
        : S(S_), R(R_), E(env_of(S_, R_)), slot(slot_), sc(sc_) {}

    // OnBackgroundHit (PathTracer.cs:384-400), OnLightHit (:402-421) and the
    // survival decision (:277-282, ApproxThroughput == PrefixWeight).
    __device__ __forceinline__ void on_hit(float4 h) {
        // decisions of this vertex (the single-kernel tracer reuses the object along the path)
        q = 1.0f;
        cont = false;
        do_nee = false;
        surface_hit = false;
        const int li = __float_as_int(h.w);
        const float nsr = (float)R.num_shadow_rays;
        if (li < 0) {
            if (depth >= R.min_depth) {
                float mis = 0.0f;
                f3 contrib = gray(0.0f);
                if (R.has_env && R.enable_bsdf_di) {
                    mis = 1.0f;
                    if (depth > 1) {
                        float pdf_ne = env_pdf(E, dir) * nsr;
                        mis = srcp(1.0f + sdiv(pdf_ne, prev_pdf));
                    }
                    contrib = env_radiance(E, dir);
                }
                L = L + prefix * mis * contrib;
                if (log) {
                    log->hit_background = true;
                    log->hit_emission = contrib;
                    log->hit_mis = mis;
                }
            }
            return;
        }
        s = surface(S, li, h.y, h.z);
        make_shader(S, s, -dir, &sh);
        surface_hit = true;
        if (ADJ) vtex = vertex_texels(S, sh, s.uv);
        int ei = S.tri_emitter[li];
        if (ei >= 0 && depth >= R.min_depth) {
            DevEmitter e = S.emitters[R.emitter_off + ei];
            float mis = 1.0f;
            if (depth > 1) {
                float jac = area_to_solid_angle(prev_pos, s.pos, s.ng);
                float pdf_ne = sdiv(sdiv(e.inv_area, (float)R.num_emitters) * nsr, jac);
                float ratio = sdiv(pdf_ne, prev_pdf);
                mis = srcp(ratio + 1.0f);
                if (!R.enable_bsdf_di) mis = 0.0f;
            }
            f3 em = emitted(e, s.ns, -dir);
            L = L + prefix * mis * em;
            if (log) {
                log->hit_light = true;
                log->hit_emission = em;
                log->hit_mis = mis;
            }
        }
        q = depth > 4 ? clampf(average(prefix), 0.05f, 0.95f) : 1.0f;
        bool stop = rng.next_float() > q || depth == R.max_depth;
        do_nee = !stop && depth + 1 >= R.min_depth;
        cont = !stop;
    }

    // PerformBackgroundNextEvent (PathTracer.cs:423-452): one environment
    // sample. Returns whether it yields a shadow ray with a non-zero
    // contribution (a zero one cannot change the estimate and is not traced).
    __device__ __forceinline__ bool env_next_event(float4* o, float4* d, float4* c, float4* g) {
        const float nsr = (float)R.num_shadow_rays;
        EnvSample es = env_sample(E, rng.next_float2());
        f3 bc;
        float pdf_bsdf;
        eval_cos_pdf(sh, es.dir, &bc, &pdf_bsdf);
        if (pdf_bsdf == 0.0f || es.pdf == 0.0f) return false;
        float mis = R.enable_bsdf_di ? srcp(1.0f + sdiv(pdf_bsdf, es.pdf * nsr)) : 1.0f;
        f3 contrib = es.weight * bc / nsr;
        f3 val = prefix * (mis * contrib) / q;
        if (log) {
            log->env_ok = true;
            log->env_nonzero = !is_zero(contrib);
            log->env_dir = es.dir;
            log->env_emission = es.weight * es.pdf;
            log->env_pdf = es.pdf;
            log->env_bc = bc;
            log->env_mis = mis;
            log->env_o = f4_of(offset_point(s, es.dir), s.err);
            log->env_d = f4_of(es.dir, INFINITY);
        }
        if (is_zero(val)) return false;
        *o = f4_of(offset_point(s, es.dir), s.err);
        *d = f4_of(es.dir, INFINITY);
        *c = f4_of(val, __int_as_float((int)shadow_tag(slot, sc, 0)));
        if (ADJ) *g = f4_of(sh.f.to_local(es.dir), 0.0f);  // the direction k_shade_grad differentiates at
        return true;
    }

    // PerformNextEventEstimation (PathTracer.cs:454-501): one area-light sample.
    __device__ __forceinline__ bool area_next_event(float4* o, float4* d, float4* c, float4* g) {
        const float nsr = (float)R.num_shadow_rays;
        int idx = (int)rng.next_int((uint32_t)R.num_emitters);
        DevEmitter e;
        float sel = srcp((float)R.num_emitters);
        f2 b = to_uniform_triangle(rng.next_float2());
        Surf lp;
        if (em_tab) {
            const float4* r = em_tab + kEmitterWords * idx;
            e = *reinterpret_cast<const DevEmitter*>(r);
            lp = surface_at(r + 2, r + 5, b.x, b.y);
        } else {
            e = S.emitters[R.emitter_off + idx];
            lp = surface(S, e.tri, b.x, b.y);
        }
        f3 l2s = normalize(s.pos - lp.pos);
        f3 em = emitted(e, lp.ns, l2s);
        float jac = area_to_solid_angle(s.pos, lp.pos, lp.ng);
        f3 bc;
        float pdf_bsdf_sa;
        eval_cos_pdf(sh, -l2s, &bc, &pdf_bsdf_sa);
        float pdf_ne = e.inv_area * sel * nsr;
        float pdf_bsdf = pdf_bsdf_sa * jac;
        f3 val = gray(0.0f);
        if (jac != 0.0f) {
            float ratio = sdiv(pdf_bsdf, pdf_ne);
            float mis = R.enable_bsdf_di ? srcp(ratio + 1.0f) : 1.0f;
            float pdf = sdiv(e.inv_area, jac) * sel * nsr;
            f3 contrib = em / pdf * bc;
            val = prefix * (mis * contrib) / q;
            if (log) {
                log->area_ok = true;
                log->area_nonzero = !is_zero(contrib);
                log->area_pos = lp.pos;
                log->area_emission = em;
                log->area_pdf = pdf;
                log->area_bc = bc;
                log->area_mis = mis;
            }
        }
        if (!log && is_zero(val)) return false;
        f3 dd = lp.pos - s.pos;
        f3 p0 = offset_point(s, dd);
        f3 p1 = offset_point(lp, -dd);
        if (log) {
            log->area_o = f4_of(p0, 1e-5f);
            log->area_d = f4_of(p1 - p0, 1.0f - 1e-5f);
            if (is_zero(val)) return false;
        }
        if (ADJ) *g = f4_of(sh.f.to_local(-l2s), 0.0f);  // the direction k_shade_grad differentiates at
        *o = f4_of(p0, 1e-5f);
        *d = f4_of(p1 - p0, 1.0f - 1e-5f);
        *c = f4_of(val, __int_as_float((int)shadow_tag(slot, sc, 1)));
        return true;
    }

    // SampleDirection (PathTracer.cs:526-530) and the path-state update
    // (:367-378). Returns false if the path ends here.
    __device__ __forceinline__ bool sample_continuation(float4* no, float4* nd) {
        float u_comp = rng.next_float();
        f2 u_dir = rng.next_float2();
        BsdfSample bs = sample_bsdf(sh, u_comp, u_dir);
        if (bs.pdf == 0.0f || is_zero(bs.weight)) return false;
        *no = f4_of(offset_point(s, bs.dir), s.err);
        *nd = f4_of(bs.dir, INFINITY);
        next_dir = bs.dir;
        if (log) log->scatter = bs.weight / q;
        prefix = prefix * (bs.weight / q);
        prev_pdf = bs.pdf * q;
        prev_pos = s.pos;
        depth = depth + 1;
        return true;
    }
};

// A world's only scene record, read through the scalar data cache (for worlds
// of one scene, where every path has scene 0).
__device__ __forceinline__ void load_rec_uniform(const SceneRec* src, SceneRec* dst) {
    typedef const __attribute__((address_space(4))) uint32_t ConstU32;
    ConstU32* q = (ConstU32*)src;
    uint32_t* d = reinterpret_cast<uint32_t*>(dst);
#pragma unroll
    for (int k = 0; k < (int)(sizeof(SceneRec) / 4); ++k) d[k] = q[k];
}

// One bounce of EstimateIncidentRadiance for every path of the queue. UNI: every
// record agrees on what this kernel reads (a world of one scene, or a view set
// rendered with one setting; DevScene::recs_alike), so record 0's fields are
// read once through the scalar data cache instead of by per-lane loads (shade
// 26.5 -> 24.8 ms per large-scene render), and a scene of at most
// kLdsEmitters emitters has them staged in LDS for the area-light samples.
// Shadow rays keep their path's own record index in their tags.
template <bool ADJ, int SB, bool UNI>
__global__ void __launch_bounds__(SB) __attribute__((amdgpu_waves_per_eu(ADJ ? SSPT_SHADE_ADJ_WAVES : SSPT_SHADE_WAVES)))
k_shade(DevScene S, Wave w, int buf, int max_nsr) {
    __shared__ uint32_t app[3 * (SB / 64) + 2];  // block_append_word scratch
    SceneRec rec_u;
    if (UNI) load_rec_uniform(S.recs, &rec_u);
    __shared__ float4 em_lds[UNI ? kLdsEmitters * kEmitterWords : 1];
    const uint32_t count = queue_count(w, buf);
    // blocks beyond the queue's length exit at once and stage nothing
    const bool em_staged = UNI && rec_u.num_emitters <= kLdsEmitters && blockIdx.x * blockDim.x < count;
    if (em_staged) {
        for (int t = threadIdx.x; t < rec_u.num_emitters * kEmitterWords; t += blockDim.x) {
            const int e = t / kEmitterWords, k = t - e * kEmitterWords;
            const DevEmitter* E = S.emitters + rec_u.emitter_off + e;
            em_lds[t] = k < 2   ? reinterpret_cast<const float4*>(E)[k]
                        : k < 5 ? S.tri_pos[3 * E->tri + (k - 2)]
                                : S.tri_aux[4 * E->tri + (k - 5)];
        }
        __syncthreads();
    }
    const int nb = buf ^ 1;
    for (uint32_t base = blockIdx.x * blockDim.x; base < count; base += gridDim.x * blockDim.x) {
        uint32_t i = base + threadIdx.x;
        bool active = i < count;
        uint32_t qs = active ? w.q_slot[buf][i] : 0u;
        uint32_t slot = qs & kSlotMask, sc = qs >> kSlotBits;
        const SceneRec& R = UNI ? rec_u : S.recs[sc];
        PathVertex<ADJ> v(S, R, slot, sc);
        if (em_staged) v.em_tab = em_lds;
        if (active) {
            v.dir = f3_of(w.q_d[buf][i]);
            v.rng.state = w.q_rng[buf][i];
            float4 thr = w.q_thr[buf][i];
            v.prefix = f3_of(thr);
            v.prev_pdf = thr.w;
            float4 pv = w.q_prev[buf][i];
            v.depth = __float_as_int(pv.w);
            v.prev_pos = f3_of(pv);
            v.L = f3_of(w.q_L[buf][i]);
            if (ADJ) {
                // the previous vertex's next-event rays have resolved: snapshot A_{depth-1}
                if (v.depth >= 2) {
                    f3 A = v.L + f3_of(w.st_Le[slot]) + f3_of(w.st_La[slot]);
                    w.ad_A[(size_t)(v.depth - 2) * w.capacity + slot] = f4_of(A, 1.0f);
                }
            }
            v.on_hit(w.hits[i]);
            // DenoiseBuffers.LogPrimaryHit (PathTracer.cs:336-339): GetScatterStrength and the shading normal
            if (R.enable_denoiser && v.depth == 1 && v.surface_hit) {
                w.st_alb[slot] = f4_of(v.sh.base, 0.0f);
                w.st_nrm[slot] = f4_of(v.s.ns, 0.0f);
            }
        }

        // Next-event estimation (PathTracer.cs:356-364): per shadow ray,
        // environment then area light. The loop bound is the render's largest
        // NumShadowRays, uniform over the block as block_append_word needs (lanes of
        // scenes with fewer shadow rays simply do not take part).
        const int nsr_i = R.num_shadow_rays;
        float4 no = make_float4(0.0f, 0.0f, 0.0f, 0.0f), nd = no;
        bool cont = false;
        uint32_t n_idx = 0;
        bool cont_appended = false;  // the same in every thread (max_nsr is a kernel argument)
        for (int k = 0; k < max_nsr; ++k) {
            const bool nee_k = v.do_nee && k < nsr_i;
            float4 eo, ed, ec, eg = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            const bool env_ray = nee_k && R.has_env && v.env_next_event(&eo, &ed, &ec, &eg);
            wave_count(w.stats, kStatShadow, nee_k && R.has_env);
            float4 ao, ad, ac, ag = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            const bool area_ray = nee_k && R.num_emitters > 0 && v.area_next_event(&ao, &ad, &ac, &ag);
            wave_count(w.stats, kStatShadow, nee_k && R.num_emitters > 0);
            // one block pass for both: the block's environment rays, then its area-light
            // rays; the forward tracer's last pass also takes the continuation ray (its
            // BSDF sample draws after the next-event samples, as in the reference)
            uint32_t e_idx, a_idx;
            if (!ADJ && k == max_nsr - 1) {
                cont = v.cont && v.sample_continuation(&no, &nd);
                const uint3 idx = block_append_word<true, true>(queue_word(w, nb), 32, env_ray, area_ray, cont, app);
                e_idx = idx.x;
                a_idx = idx.y;
                n_idx = idx.z;
                cont_appended = true;
            } else {
                const uint3 idx = block_append_word<true, false>(queue_word(w, nb), 32, env_ray, area_ray, false, app);
                e_idx = idx.x;
                a_idx = idx.y;
            }
            if (env_ray) {
                w.s_o[e_idx] = eo;
                w.s_d[e_idx] = ed;
                w.s_c[e_idx] = ec;
                if (ADJ) {  // the local direction and the queue entry, for k_shade_grad
                    w.s_g[e_idx] = make_float4(eg.x, eg.y, eg.z, __int_as_float((int)i));
                    w.s_t[e_idx] = v.vtex;
                }
            }
            if (area_ray) {
                w.s_o[a_idx] = ao;
                w.s_d[a_idx] = ad;
                w.s_c[a_idx] = ac;
                if (ADJ) {
                    w.s_g[a_idx] = make_float4(ag.x, ag.y, ag.z, __int_as_float((int)i));
                    w.s_t[a_idx] = v.vtex;
                }
            }
        }

        if (!cont_appended) cont = v.cont && v.sample_continuation(&no, &nd);
        if (ADJ && active) {
            // the vertex's record (index depth - 1): its texels and, for k_shade_grad, the
            // sampled direction whose d log f (the sampling pdf detached) it computes;
            // (-1, -1) where the path ends here
            const int vd = cont ? v.depth - 1 : v.depth;
            if (vd >= 1) {
                const size_t rec = (size_t)(vd - 1) * w.capacity + slot;
                w.ad_tex[rec] = cont ? v.vtex : make_int2(-1, -1);
                if (cont && (v.vtex.x >= 0 || v.vtex.y >= 0)) w.ad_dlog[rec * 3] = f4_of(v.next_dir, 0.0f);
            }
        }
        if (!cont_appended) n_idx = block_append_word<false, false>(queue_word(w, nb), 0, cont, false, false, app).x;
        if (cont) {
            w.q_o[nb][n_idx] = no;
            w.q_d[nb][n_idx] = nd;
            w.q_slot[nb][n_idx] = qs;
            w.q_rng[nb][n_idx] = v.rng.state;
            w.q_thr[nb][n_idx] = f4_of(v.prefix, v.prev_pdf);
            w.q_prev[nb][n_idx] = f4_of(v.prev_pos, __int_as_float(v.depth));
            w.q_L[nb][n_idx] = f4_of(v.L, 0.0f);
        } else if (active) {
            // the path ends here: its radiance from hits and its last depth, by slot
            w.st_L[slot] = f4_of(v.L, 0.0f);
            w.st_prev[slot] = make_float4(0.0f, 0.0f, 0.0f, __int_as_float(v.depth));
        }
    }
}

// The adjoint's material derivatives of one bounce, apart from the adjoint
// shade so that their dual-number evaluations do not crowd its registers. Jobs
// [0, queue length): the continuation of each textured vertex, d log f of its
// sampled direction into the vertex's record; the rest: each next-event ray of
// a textured vertex, the gradient of its contribution (val * delta * d log f,
// summed over the colour channels) into s_g, which held the ray's local
// direction and queue entry. The vertex's shading context is rebuilt from the
// queue's hit and direction, so this runs after the bounce's shade and before
// the next trace overwrites the hits.
__global__ void __launch_bounds__(256) k_shade_grad(DevScene S, Wave w, int buf) {
    const uint32_t qn = queue_count(w, buf), sn = shadow_count(w, buf ^ 1);
    for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < qn + sn; j += gridDim.x * blockDim.x) {
        const bool shadow = j >= qn;
        uint32_t i, r = 0;
        size_t rec = 0;
        int2 tex;
        float4 aux;
        if (!shadow) {
            i = j;
            const uint32_t slot = w.q_slot[buf][i] & kSlotMask;
            const int depth = __float_as_int(w.q_prev[buf][i].w);
            rec = (size_t)(depth - 1) * w.capacity + slot;
            tex = w.ad_tex[rec];
            if (tex.x < 0 && tex.y < 0) continue;
            aux = w.ad_dlog[rec * 3];
        } else {
            r = j - qn;
            tex = w.s_t[r];
            if (tex.x < 0 && tex.y < 0) continue;
            aux = w.s_g[r];
            i = (uint32_t)__float_as_int(aux.w);
        }
        const float4 h = w.hits[i];
        const Surf s = surface(S, __float_as_int(h.w), h.y, h.z);
        Shader sh;
        make_shader(S, s, -f3_of(w.q_d[buf][i]), &sh);
        const f3 in = shadow ? f3_of(aux) : sh.f.to_local(f3_of(aux));
        const D3 f = material_value_dual(S, sh, s.uv, in, tex.x >= 0, tex.y >= 0);
        float dl[12];
        dlog_of(f, dl);
        if (!shadow) {
            w.ad_dlog[rec * 3 + 0] = make_float4(dl[0], dl[1], dl[2], dl[3]);
            w.ad_dlog[rec * 3 + 1] = make_float4(dl[4], dl[5], dl[6], dl[7]);
            w.ad_dlog[rec * 3 + 2] = make_float4(dl[8], dl[9], dl[10], dl[11]);
        } else {
            const float4 c = w.s_c[r];
            const uint32_t tag = (uint32_t)__float_as_int(c.w);
            const SceneRec& R = S.recs[tag >> (kSlotBits + 1)];
            const uint32_t slot = (tag >> 1) & kSlotMask;
            const float* dp = S.adj_image + R.fb_off + 3ull * (w.st_pix[slot] & kSlotMask);
            const f3 delta = mk3(dp[0], dp[1], dp[2]) * R.adj_scale;
            float g[4];
#pragma unroll
            for (int k = 0; k < 4; ++k)
                g[k] = delta.x * c.x * dl[k] + delta.y * c.y * dl[4 + k] + delta.z * c.z * dl[8 + k];
            w.s_g[r] = make_float4(g[0], g[1], g[2], g[3]);
        }
    }
}

// The plain path tracer: one thread follows one path from its camera ray to
// its end (RenderPixel + EstimateIncidentRadiance in a single kernel, shadow
// rays traced where they are sampled). Same arithmetic and RNG draws as the
// wavefront, so both render the same image; it exists as the per-path-launch
// baseline the wavefront is measured against. Writes the path's radiance sums
// to the slot for the accumulate kernels.
__global__ void __launch_bounds__(kTraceBlock) k_path(DevScene S, Wave w) {
    __shared__ int stack_node[kStackNodeLds];
    __shared__ float stack_t[kStackLds * kTraceBlock];
    const WaveDesc d = cur_wave(w);
    for (uint32_t slot = blockIdx.x * blockDim.x + threadIdx.x; slot < d.count; slot += gridDim.x * blockDim.x) {
        PathStart p = start_path(S, d.start + slot);
        const SceneRec& R = S.recs[p.scene];
        PathVertex<false> v(S, R, slot, (uint32_t)p.scene);
        v.rng = p.rng;
        v.prefix = gray(1.0f);
        v.depth = 1;
        f3 Le = gray(0.0f), La = gray(0.0f);
        f3 alb = gray(0.0f), nrm = gray(0.0f);
        float4 o4 = f4_of(p.o, 0.0f), d4 = f4_of(p.dir, INFINITY);
        for (;;) {
            HitOut h = traverse<false>(S, R.root, R.abs_max, f3_of(o4), f3_of(d4), o4.w, d4.w,
                                       stack_node + threadIdx.x, stack_t + threadIdx.x);
            wave_count(w.stats, kStatRays, true);
            wave_count(w.stats, kStatHits, h.li >= 0);
            v.dir = f3_of(d4);
            v.on_hit(h.li >= 0 ? make_float4(h.t, h.u, h.v, __int_as_float(h.li))
                               : make_float4(INFINITY, 0.0f, 0.0f, __int_as_float(-1)));
            if (v.depth == 1 && v.surface_hit) {
                alb = v.sh.base;
                nrm = v.s.ns;
            }
            for (int k = 0; v.do_nee && k < R.num_shadow_rays; ++k) {
                float4 so, sd, c, g;
                if (R.has_env) {
                    wave_count(w.stats, kStatShadow, true);
                    if (v.env_next_event(&so, &sd, &c, &g)) {
                        bool occ = traverse<true>(S, R.root, R.abs_max, f3_of(so), f3_of(sd), so.w, sd.w,
                                                  stack_node + threadIdx.x, stack_t + threadIdx.x)
                                       .li >= 0;
                        wave_count(w.stats, kStatShadowTraced, true);
                        wave_count(w.stats, kStatOccluded, occ);
                        if (!occ) Le = Le + f3_of(c);
                    }
                }
                if (R.num_emitters > 0) {
                    wave_count(w.stats, kStatShadow, true);
                    if (v.area_next_event(&so, &sd, &c, &g)) {
                        bool occ = traverse<true>(S, R.root, R.abs_max, f3_of(so), f3_of(sd), so.w, sd.w,
                                                  stack_node + threadIdx.x, stack_t + threadIdx.x)
                                       .li >= 0;
                        wave_count(w.stats, kStatShadowTraced, true);
                        wave_count(w.stats, kStatOccluded, occ);
                        if (!occ) La = La + f3_of(c);
                    }
                }
            }
            if (!v.cont || !v.sample_continuation(&o4, &d4)) break;
        }
        w.st_pix[slot] = p.pix | ((uint32_t)p.scene << kSlotBits);
        w.st_prev[slot] = make_float4(0.0f, 0.0f, 0.0f, __int_as_float(v.depth));
        w.st_L[slot] = f4_of(v.L, 0.0f);
        w.st_Le[slot] = f4_of(Le, 0.0f);
        w.st_La[slot] = f4_of(La, 0.0f);
        if (R.enable_denoiser) {
            w.st_alb[slot] = f4_of(alb, 0.0f);
            w.st_nrm[slot] = f4_of(nrm, 0.0f);
        }
    }
}

// ReplayPixel (PathTracer.cs:61-69): thread 0 re-traces one path exactly as
// k_path traces it (same vertex object, same RNG draws, same shadow rays) and
// writes the PathGraph nodes RenderPixel / EstimateIncidentRadiance add
// (PathTracer.cs:303, 326, 346-348, 447, 496). The block has kTraceBlock
// threads only because the traversal stack's LDS layout is strided by it.
__device__ __forceinline__ PathNode* replay_node(PathNode* nodes, uint32_t cap, uint32_t* n, PathNode* spare, int kind,
                                                 int ancestor, f3 pos) {
    PathNode* x = *n < cap ? nodes + *n : spare;
    *n += 1;
    *x = PathNode{};
    x->kind = kind;
    x->ancestor = ancestor;
    x->position[0] = pos.x;
    x->position[1] = pos.y;
    x->position[2] = pos.z;
    return x;
}
__device__ __forceinline__ void put3(float* dst, f3 v) {
    dst[0] = v.x;
    dst[1] = v.y;
    dst[2] = v.z;
}

__global__ void __launch_bounds__(kTraceBlock) k_replay(DevScene S, int scene, uint32_t pix, uint32_t iteration,
                                                        PathNode* nodes, uint32_t cap, uint32_t* count,
                                                        float* estimate) {
    __shared__ int stack_node[kStackNodeLds];
    __shared__ float stack_t[kStackLds * kTraceBlock];
    if (threadIdx.x != 0) return;
    const SceneRec& R = S.recs[scene];
    const uint32_t row = pix / (uint32_t)R.cam.width, col = pix - row * (uint32_t)R.cam.width;
    PathVertex<false> v(S, R, 0, (uint32_t)scene);
    ReplayLog log;
    v.log = &log;
    v.rng = Rng{hash_seed(R.base_seed, pix, iteration)};
    f2 off = v.rng.next_float2();
    f3 o, d;
    camera_ray(R.cam, (float)col + off.x, (float)row + off.y, &o, &d);
    v.prefix = gray(1.0f);
    v.depth = 1;
    uint32_t n = 0;
    PathNode spare;
    replay_node(nodes, cap, &n, &spare, 0, -1, o);
    int cur = 0;
    f3 cur_pos = o, scatter = gray(1.0f);
    float survival = 1.0f;
    f3 Le = gray(0.0f), La = gray(0.0f);
    float4 o4 = f4_of(o, 0.0f), d4 = f4_of(d, INFINITY);
    int* sn = stack_node + threadIdx.x;
    float* st = stack_t + threadIdx.x;
    for (;;) {
        HitOut h = traverse<false>(S, R.root, R.abs_max, f3_of(o4), f3_of(d4), o4.w, d4.w, sn, st);
        v.dir = f3_of(d4);
        log = ReplayLog{};
        v.on_hit(h.li >= 0 ? make_float4(h.t, h.u, h.v, __int_as_float(h.li))
                           : make_float4(INFINITY, 0.0f, 0.0f, __int_as_float(-1)));
        if (!v.surface_hit) {
            if (log.hit_background) {
                PathNode* x = replay_node(nodes, cap, &n, &spare, 3, cur, cur_pos + v.dir);
                put3(x->emission, log.hit_emission);
                x->mis_weight = log.hit_mis;
            }
            break;
        }
        PathNode* x = replay_node(nodes, cap, &n, &spare, 1, cur, v.s.pos);
        put3(x->scatter_weight, scatter);
        x->survival_probability = survival;
        if (log.hit_light) {
            put3(x->emission, log.hit_emission);
            x->mis_weight = log.hit_mis;
            x->flags = 1u;
        }
        cur = (int)n - 1;
        cur_pos = v.s.pos;
        for (int k = 0; v.do_nee && k < R.num_shadow_rays; ++k) {
            float4 so, sd, c, g;
            if (R.has_env) {
                log.env_ok = false;
                bool traced = v.env_next_event(&so, &sd, &c, &g);
                if (log.env_ok && log.env_nonzero) {
                    bool occ = traverse<true>(S, R.root, R.abs_max, f3_of(log.env_o), f3_of(log.env_d), log.env_o.w,
                                              log.env_d.w, sn, st)
                                   .li >= 0;
                    if (!occ) {
                        PathNode* y = replay_node(nodes, cap, &n, &spare, 2, cur, cur_pos + log.env_dir);
                        put3(y->emission, log.env_emission);
                        y->pdf = log.env_pdf;
                        put3(y->bsdf_cos, log.env_bc);
                        y->mis_weight = log.env_mis;
                        put3(y->prefix_weight, v.prefix);
                        y->flags = 2u;
                        if (traced) Le = Le + f3_of(c);
                    }
                }
            }
            if (R.num_emitters > 0) {
                log.area_ok = false;
                bool traced = v.area_next_event(&so, &sd, &c, &g);
                if (log.area_ok && log.area_nonzero) {
                    bool occ = traverse<true>(S, R.root, R.abs_max, f3_of(log.area_o), f3_of(log.area_d),
                                              log.area_o.w, log.area_d.w, sn, st)
                                   .li >= 0;
                    if (!occ) {
                        PathNode* y = replay_node(nodes, cap, &n, &spare, 2, cur, log.area_pos);
                        put3(y->emission, log.area_emission);
                        y->pdf = log.area_pdf;
                        put3(y->bsdf_cos, log.area_bc);
                        y->mis_weight = log.area_mis;
                        put3(y->prefix_weight, v.prefix);
                        if (traced) La = La + f3_of(c);
                    }
                }
            }
        }
        if (!v.cont || !v.sample_continuation(&o4, &d4)) break;
        scatter = log.scatter;
        survival = v.q;
    }
    *count = n;
    // the order k_accumulate_* sums a path's parts in
    f3 total = v.L + Le + La;
    estimate[0] = total.x;
    estimate[1] = total.y;
    estimate[2] = total.z;
}

// DebugVisualizer.Render / RenderPixel / ComputeColor
// (SeeSharp/Integrators/DebugVisualizer.cs:20-59): per pixel and iteration the
// jittered primary ray's hit shaded by |cos| between the face normal and the
// direction back to the camera ("eye light"), 0 on a miss, splatted into the
// framebuffer's running mean. One thread per pixel walks the iterations in order.
__global__ void __launch_bounds__(kTraceBlock) k_debug_visualize(DevScene S, int scene, uint32_t base_seed,
                                                                 uint32_t first, uint32_t num, float* image) {
    __shared__ int stack_node[kStackNodeLds];
    __shared__ float stack_t[kStackLds * kTraceBlock];
    const SceneRec& R = S.recs[scene];
    for (uint32_t pix = blockIdx.x * blockDim.x + threadIdx.x; pix < R.num_pixels; pix += gridDim.x * blockDim.x) {
        const uint32_t row = pix / (uint32_t)R.cam.width, col = pix - row * (uint32_t)R.cam.width;
        f3 img = mk3(image[3 * pix], image[3 * pix + 1], image[3 * pix + 2]);
        for (uint32_t it = first; it < first + num; ++it) {
            Rng rng{hash_seed(base_seed, pix, it)};
            f2 off = rng.next_float2();
            f3 o, d;
            camera_ray(R.cam, (float)col + off.x, (float)row + off.y, &o, &d);
            HitOut h = traverse<false>(S, R.root, R.abs_max, o, d, 0.0f, INFINITY, stack_node + threadIdx.x,
                                       stack_t + threadIdx.x);
            float value = 0.0f;
            if (h.li >= 0) {
                const f3 from = -d;
                const f3 n = surface(S, h.li, h.u, h.v).ng;
                value = fabsf(dot(n, from));
                value /= length(n);
                value /= length(from);
            }
            const float cnt = (float)(it + 1);
            if (it + 1 > 1) img = img * ((cnt - 1.0f) / cnt);
            img = img + div_exact(gray(value), cnt);
        }
        image[3 * pix] = img.x;
        image[3 * pix + 1] = img.y;
        image[3 * pix + 2] = img.z;
    }
}

// ------------------------------------------------------------ light tracer -
// PerspectiveCamera.SampleResponse (PerspectiveCamera.cs:111-146) of the pinhole:
// the pixel that sees `pos` (WorldToFilm, :148-166) and the importance weight
// SolidAngleToPixelJacobian(pos) (:177-191) * Camera.SurfaceAreaToSolidAngleJacobian
// (Camera.cs:91-98) of the surface normal. Correctly rounded divisions and
// square roots, as for camera rays, so the film position picks the CPU's pixel.
__device__ __forceinline__ bool camera_response(const DevCamera& c, f3 pos, f3 normal, uint32_t* pix, float* weight) {
    const float* M = c.w2c;
    const f3 l = mk3(pos.x * M[0] + pos.y * M[4] + pos.z * M[8] + M[12],
                     pos.x * M[1] + pos.y * M[5] + pos.z * M[9] + M[13],
                     pos.x * M[2] + pos.y * M[6] + pos.z * M[10] + M[14]);
    if (l.z > 0.0f) return false;
    const float* P = c.c2v;  // Vector4.Transform(local, cameraToView), w = 1
    const float vx = l.x * P[0] + l.y * P[4] + l.z * P[8] + P[12];
    const float vy = l.x * P[1] + l.y * P[5] + l.z * P[9] + P[13];
    const float vw = l.x * P[3] + l.y * P[7] + l.z * P[11] + P[15];
    const float fx = (vx / vw + 1.0f) / 2.0f * (float)c.width;
    const float fy = (1.0f - vy / vw) / 2.0f * (float)c.height;
    if (fx < 0.0f || fx > (float)c.width || fy < 0.0f || fy > (float)c.height) return false;
    const float cosine = l.z / sqrtf(dot(l, l));
    const float d = c.image_plane / cosine;
    float jac = d * d / fabsf(cosine);
    const f3 to_cam = mk3(c.pos[0], c.pos[1], c.pos[2]) - pos;
    const float dist = sqrtf(dot(to_cam, to_cam));
    const float cos_cam = fabsf(dot(normal, to_cam)) / dist;
    jac *= dist == 0.0f ? 0.0f : cos_cam / (dist * dist);
    const int px = min(max(to_int(fx), 0), c.width - 1), py = min(max(to_int(fy), 0), c.height - 1);
    *pix = (uint32_t)py * (uint32_t)c.width + (uint32_t)px;
    *weight = jac;
    return true;
}

struct EmitterRay {
    Surf p;
    f3 dir;
    float pdf;
    f3 weight;
};
// Emitter.SampleRay: DiffuseEmitter (DiffuseEmitter.cs:44-64) and GlossyEmitter
// (GlossyEmitter.cs:47-70, SampleWarp.ToCosineLobe SampleWarp.cs:171-182) from a
// point of Triangle.SampleUniformArea (Triangle.cs:42-56).
__device__ __forceinline__ EmitterRay emitter_ray(const DevScene& S, const DevEmitter& e, f2 primary_pos,
                                                  f2 primary_dir) {
    EmitterRay r;
    const f2 b = to_uniform_triangle(primary_pos);
    r.p = surface(S, e.tri, b.x, b.y);
    const float pos_pdf = e.inv_area;
    Frame f;
    f.build(r.p.ns);
    f3 local;
    float pdf;
    const f3 radiance = mk3(e.r, e.g, e.b);
    if (e.kind == kEmitterGlossy) {
        const float power = e.exponent + 1.0f;
        const float phi = kPi * 2.0f * primary_dir.x;
        const float cos_t = powf(primary_dir.y, srcp(power + 1.0f));
        const float sin_t = ssqrt(1.0f - cos_t * cos_t);
        float sp, cp;
        ssincos(phi, &sp, &cp);
        local = mk3(sin_t * cp, sin_t * sp, cos_t);
        pdf = sdiv((power + 1.0f) * powf(cos_t, power) * 1.0f, 2.0f * kPi);
        r.weight = radiance * powf(cos_t, e.exponent + 1.0f) * e.norm / pos_pdf / pdf;
    } else {
        local = to_cos_hemisphere(primary_dir);
        pdf = sdiv(local.z, kPi);
        r.weight = radiance / pos_pdf * kPi;
    }
    r.dir = f.to_world(local);
    r.pdf = pdf * pos_pdf;
    return r;
}

// SurfaceShader.Evaluate and Pdf(inDir) of one direction in one pass: the BSDF
// value without the cosine, the forward and the reverse pdf (the bidirectional
// weights need all three; two calls would evaluate the material twice).
__device__ __forceinline__ f3 value_pdf_pair(const Shader& sh, f3 in_world, float* fwd, float* rev) {
    f3 v;
    if (sh.m.kind == 1) {
        const f3 in = sh.f.to_local(in_world);
        generic_value_pdf(sh, in, &v, fwd);
        *rev = generic_pdf_rev(sh, in);
    } else {
        diffuse_value_pdf(sh, in_world, &v, fwd);
        *rev = diffuse_pdf_rev(sh, in_world);
    }
    return v;
}
// The same with EvaluateWithCosine's |cos| of the shading frame.
__device__ __forceinline__ f3 value_cos_pdf_pair(const Shader& sh, f3 in_world, float* fwd, float* rev) {
    const f3 v = value_pdf_pair(sh, in_world, fwd, rev);
    return v * fabsf(sh.f.to_local(in_world).z);
}

// SurfaceShader.Evaluate: the BSDF value without the cosine.
__device__ __forceinline__ f3 material_value(const Shader& sh, f3 in_world) {
    f3 v;
    float pdf;
    if (sh.m.kind == 1) generic_value_pdf(sh, sh.f.to_local(in_world), &v, &pdf);
    else diffuse_value_pdf(sh, in_world, &v, &pdf);
    return v;
}

// BidirBase.TraceLightPaths + SplatLightVertices for one iteration of a
// VertexCacheBidir set up as a pure light tracer: one thread follows one light
// path (SelectLight, TraceEmitterPath, BidirBase.Light.cs:103-110,212-277;
// RandomWalk.StartFromEmitter / ContinueWalk, RandomWalk.cs:82-90,150-217) and
// connects each vertex to the camera as the walk stores it
// (ConnectLightVertexToCamera, BidirBase.Light.cs:35-92, MIS weight 1): the
// connection draws no random numbers, so this equals splatting the cached
// vertices afterwards. Splats are added to the running-mean image with fp32
// atomics (FrameBuffer.Splat: value / CurIteration; non-finite ones counted).
__global__ void __launch_bounds__(kTraceBlock, 4) k_light_paths(DevScene S, int scene, LightParams P, float* image,
                                                             unsigned long long* stats) {
    __shared__ int stack_node[kStackNodeLds];
    __shared__ float stack_t[kStackLds * kTraceBlock];
    const SceneRec& R = S.recs[scene];
    const f3 cam = mk3(R.cam.pos[0], R.cam.pos[1], R.cam.pos[2]);
    for (uint32_t idx = blockIdx.x * blockDim.x + threadIdx.x; idx < P.num_paths; idx += gridDim.x * blockDim.x) {
        Rng rng{hash_seed(P.base_seed, idx, P.iteration)};
        // SelectLight: no background, so BackgroundProbability = 0 draws nothing
        const DevEmitter e = S.emitters[R.emitter_off + (int)rng.next_int((uint32_t)R.num_emitters)];
        const float sel = (1.0f - 0.0f) / (float)R.num_emitters;
        const f2 ppos = rng.next_float2();
        const f2 pdir = rng.next_float2();
        EmitterRay er = emitter_ray(S, e, ppos, pdir);
        er.pdf *= sel;
        if (er.pdf == 0.0f || is_zero(er.weight)) continue;
        f3 prefix = er.weight / sel;
        f3 approx = gray(1.0f);
        Surf prev = er.p;
        float pdf_dir = er.pdf;
        f3 o = offset_point(er.p, er.dir), d = er.dir;
        float tmin = er.p.err;
        for (int depth = 1; depth < P.max_depth;) {
            const HitOut h = traverse<false>(S, R.root, R.abs_max, o, d, tmin, INFINITY, stack_node + threadIdx.x,
                                             stack_t + threadIdx.x);
            wave_count(stats, kStatRays, true);
            wave_count(stats, kStatHits, h.li >= 0);
            if (h.li < 0) break;
            const Surf s = surface(S, h.li, h.u, h.v);
            if (pdf_dir * area_to_solid_angle(prev.pos, s.pos, s.ng) == 0.0f) break;
            uint32_t pix;
            float jac;
            if (depth + 1 >= P.min_depth && camera_response(R.cam, s.pos, s.ng, &pix, &jac)) {
                // Raytracer.IsOccluded(vertex, camera position): only the surface end is offset
                const f3 p0 = offset_point(s, cam - s.pos);
                const bool occ = traverse<true>(S, R.root, R.abs_max, p0, cam - p0, 1e-5f, 1.0f - 1e-5f,
                                                stack_node + threadIdx.x, stack_t + threadIdx.x)
                                     .li >= 0;
                wave_count(stats, kStatShadow, true);
                wave_count(stats, kStatShadowTraced, true);
                wave_count(stats, kStatOccluded, occ);
                if (!occ) {
                    f3 to_cam = cam - s.pos;
                    to_cam = to_cam / length(to_cam);
                    const f3 to_anc = normalize(prev.pos - s.pos);
                    Shader sh;
                    make_shader(S, s, to_cam, &sh);  // the camera side: IsOnLightSubpath = false
                    f3 f = material_value(sh, to_anc);
                    if (!is_zero(f)) {
                        f = f * sdiv(fabsf(dot(s.ns, to_anc)), fabsf(dot(s.ng, to_anc)));
                        const f3 w = prefix * f * jac / P.num_paths_f;
                        const bool bad = !isfinite(average(w));
                        wave_count(stats, kStatNan, bad);
                        if (!bad) {
                            const f3 add = div_exact(w, P.n);
                            float* px = image + 3ull * pix;
                            atomicAdd(px, add.x);
                            atomicAdd(px + 1, add.y);
                            atomicAdd(px + 2, add.z);
                        }
                    }
                }
            }
            if (depth + 1 >= P.max_depth) break;
            const float q = depth > 4 ? clampf(average(approx), 0.05f, 0.95f) : 1.0f;
            if (rng.next_float() > q) break;
            Shader sh;
            make_shader(S, s, -d, &sh);
            sh.light = true;
            const float pc = rng.next_float();
            const f2 pd = rng.next_float2();
            const BsdfSample bs = sample_bsdf(sh, pc, pd);
            approx = approx * (bs.weight / q);
            if (bs.pdf == 0.0f || is_zero(bs.weight)) break;
            // RandomWalk.cs:193-204: the geometric instead of the shading cosine, and the
            // rendering equation's cosine ratio at this vertex of the light subpath
            f3 w = bs.weight * sdiv(fabsf(dot(s.ng, bs.dir)), fabsf(dot(s.ns, bs.dir)));
            w = w * sdiv(fabsf(dot(s.ns, -d)), fabsf(dot(s.ng, -d)));
            prefix = prefix * (w / q);
            ++depth;
            pdf_dir = bs.pdf;
            prev = s;
            o = offset_point(s, bs.dir);
            d = bs.dir;
            tmin = s.err;
        }
    }
}

// ---------------------------------------------- bidirectional path tracing -
// ClassicBidir (ClassicBidir.cs) per iteration: k_bidir_light traces one light
// path per pixel into a cache in HBM, one level per vertex depth so that the
// lanes of a wave (neighbouring paths) touch neighbouring entries, and splats
// its vertices onto the camera; the camera wavefront (k_bc_*) then follows one
// camera path per pixel and combines its emitter hits, next-event samples and
// connections to the vertices of its pixel's light path. The MIS weights read
// the pdfs along both subpaths from the cache and the per-pixel camera pdf
// levels instead of per-thread arrays.

// Emitter.PdfRay: DiffuseEmitter.cs:39-42, GlossyEmitter.cs:42-45 (with
// SampleWarp.ToCosineLobeJacobian, SampleWarp.cs:192-194).
__device__ __forceinline__ float emitter_pdf_ray(const DevEmitter& e, f3 ns, f3 dir) {
    const float cosine = sdiv(dot(ns, dir), length(dir));
    if (e.kind == kEmitterGlossy) {
        const float power = e.exponent + 1.0f;
        const float jac = cosine > 0.0f ? sdiv((power + 1.0f) * powf(cosine, power) * 1.0f, 2.0f * kPi) : 0.0f;
        return e.inv_area * jac;
    }
    return e.inv_area * fmaxf(cosine, 0.0f) * kInvPi;
}

// PerspectiveCamera.SolidAngleToPixelJacobian (PerspectiveCamera.cs:177-191)
__device__ __forceinline__ float pixel_jacobian(const DevCamera& c, f3 pos) {
    const float* M = c.w2c;
    const f3 l = mk3(pos.x * M[0] + pos.y * M[4] + pos.z * M[8] + M[12],
                     pos.x * M[1] + pos.y * M[5] + pos.z * M[9] + M[13],
                     pos.x * M[2] + pos.y * M[6] + pos.z * M[10] + M[14]);
    const float cosine = l.z / sqrtf(dot(l, l));
    const float d = c.image_plane / cosine;
    return d * d / fabsf(cosine);
}

// BidirBase.BackgroundProbability (BidirBase.Light.cs:132-133), which is also
// ComputeNextEventBackgroundProbability (BidirBase.Camera.cs:354-355)
__device__ __forceinline__ float background_probability(const SceneRec& R) {
    return R.has_env ? 1.0f / (1.0f + (float)R.num_emitters) : 0.0f;
}

// SampleWarp.ToConcentricDisc (SampleWarp.cs:219-247): Shirley and Chiu's
// concentric map of the unit square onto the unit disc.
__device__ __forceinline__ f2 to_concentric_disc(f2 p) {
    float phi, r;
    const float a = 2.0f * p.x - 1.0f, b = 2.0f * p.y - 1.0f;
    if (a > -b) {
        if (a > b) {
            r = a;
            phi = (kPi * 0.25f) * sdiv(b, a);
        } else {
            r = b;
            phi = (kPi * 0.25f) * (2.0f - sdiv(a, b));
        }
    } else if (a < b) {
        r = -a;
        phi = (kPi * 0.25f) * (4.0f + sdiv(b, a));
    } else {
        r = -b;
        phi = b != 0.0f ? (kPi * 0.25f) * (6.0f - sdiv(a, b)) : 0.0f;
    }
    float sp, cp;
    ssincos(phi, &sp, &cp);
    return f2{r * cp, r * sp};
}

// EnvironmentMap.SampleRay (EnvironmentMap.cs:87-103): a direction of the map and
// a point on the disc of radius SceneRadius about SceneCenter + SceneRadius *
// direction, orthogonal to it; the ray runs back along the direction.
struct EnvRay {
    f3 o, d, weight;
    float pdf;
};
__device__ __forceinline__ float background_pos_pdf(const SceneRec& R) {
    return sdiv(1.0f / kPi, R.scene_radius * R.scene_radius);  // ToConcentricDiscJacobian / R^2
}
__device__ EnvRay env_sample_ray(const Env& E, const SceneRec& R, f2 primary_pos, f2 primary_dir) {
    const EnvSample ds = env_sample(E, primary_dir);
    const f2 disc = to_concentric_disc(primary_pos);
    Frame f;
    f.build(ds.dir);
    const f3 centre = mk3(R.scene_center[0], R.scene_center[1], R.scene_center[2]);
    EnvRay r;
    r.o = centre + R.scene_radius * (ds.dir + f.t * disc.x + f.b * disc.y);
    r.d = -ds.dir;
    const float pos_pdf = background_pos_pdf(R);
    r.weight = ds.weight / pos_pdf;
    r.pdf = pos_pdf * ds.pdf;
    return r;
}
// BidirBase.ComputeBackgroundPdf (BidirBase.Light.cs:174-178) over
// EnvironmentMap.RayPdf (EnvironmentMap.cs:127-130): the pdf of the background
// ray along `light_to_surface`, with the background's selection probability.
__device__ __forceinline__ float background_pdf(const Env& E, const SceneRec& R, f3 light_to_surface) {
    const float pdf = background_pos_pdf(R) * env_pdf(E, -light_to_surface);
    return pdf * background_probability(R);
}

// Light vertex k of path p in the cache: a surface point, or the ray origin
// (kept in lv_a, triangle -1) for vertex 0 of a path from the background. The
// bidirectional kernels are instantiated with and without a background (ENV),
// so scenes without one compile none of its code.
template <bool ENV>
__device__ __forceinline__ Surf light_vertex(const DevScene& S, const BidirParams& P, int k, uint32_t p) {
    const size_t at = (size_t)k * P.num_paths + p;
    const float4 b = P.lv_b[at];
    const int tri = __float_as_int(b.w);
    if (ENV && tri < 0) {
        Surf s{};
        s.pos = f3_of(P.lv_a[at]);
        return s;
    }
    return surface(S, tri, b.x, b.y);
}
// Whether light path p starts at the background (vertex 0 on no triangle).
template <bool ENV>
__device__ __forceinline__ bool from_background(const BidirParams& P, uint32_t p) {
    return ENV && __float_as_int(P.lv_b[p].w) < 0;
}

template <bool ENV>
struct BidirCtx {
    const DevScene& S;
    const SceneRec& R;
    const BidirParams& P;
    uint32_t path;  // the light path / pixel
    float bg;       // BackgroundProbability (0 without a background)
    __device__ __forceinline__ float4 va(int k) const { return P.lv_a[(size_t)k * P.num_paths + path]; }
    __device__ __forceinline__ float4 vb(int k) const { return P.lv_b[(size_t)k * P.num_paths + path]; }
    __device__ __forceinline__ float2 cam(int k) const { return P.cam[(size_t)k * P.num_pixels + path]; }
    __device__ __forceinline__ Surf vertex(int k) const { return light_vertex<ENV>(S, P, k, path); }
    // ClassicBidir.NextEventPdf (ClassicBidir.cs:42-44) over BidirBase.NextEventPdf
    // (BidirBase.Camera.cs:338-347) of a point on emitter e
    __device__ __forceinline__ float next_event_pdf(const DevEmitter& e) const {
        return e.inv_area * (1.0f / (float)R.num_emitters) * (1.0f - bg) * (float)P.num_shadow_rays;
    }
    // The same of the background point `to` seen from `from`: the map's pdf of the
    // (unnormalised) direction between them
    __device__ __forceinline__ float next_event_pdf_background(const Env& E, f3 from, f3 to) const {
        return env_pdf(E, to - from) * bg * (float)P.num_shadow_rays;
    }
    // BidirBase.ComputeEmitterPdf (BidirBase.Light.cs:160-166) with SelectLightPmf (:116-122)
    __device__ __forceinline__ float emitter_pdf(const DevEmitter& e, f3 ns, f3 light_to_surface, float jac) const {
        float pdf = emitter_pdf_ray(e, ns, light_to_surface);
        pdf *= jac;
        pdf *= (1.0f - bg) / (float)R.num_emitters;
        return pdf;
    }
    // ClassicBidir.CameraPathReciprocals (ClassicBidir.cs:164-176) over the
    // camera prefix: PdfsCameraToLight[i] = camera vertex i's PdfFromAncestor,
    // PdfsLightToCamera[i] = camera vertex i+1's PdfToAncestor, except
    // [last] = x_last and [last - 1] = x_prev (GatherCameraPdfs and the
    // technique's own values, BidirPathPdfs.cs:46-53).
    __device__ float camera_reciprocals(int last, float x_last, float x_prev) const {
        float sum = 0.0f, next = 1.0f;
        for (int i = last; i > 0; --i) {
            const float l2c = i == last ? x_last : (i == last - 1 ? x_prev : cam(i + 1).y);
            next *= sdiv(l2c, cam(i).x);
            if (P.connections) sum += next;
        }
        if (P.light_tracer) {
            const float l2c = last == 0 ? x_last : (last == 1 ? x_prev : cam(1).y);
            sum += sdiv(next * l2c, cam(0).x) * P.num_paths_f;
        }
        return sum;
    }
    // ClassicBidir.LightPathReciprocals (ClassicBidir.cs:187-198) of the full
    // path of c camera vertices and light vertices s .. 0 (N = c + s + 1 pdfs):
    // PdfsCameraToLight[c] = x0, [c + 1] = x1, [i >= c + 2] = light vertex
    // (s + c + 2 - i)'s PdfReverseAncestor; PdfsLightToCamera[i < N - 1] =
    // light vertex (s - i + c)'s PdfFromAncestor, [N - 1] = 1
    // (GatherLightPdfs, BidirPathPdfs.cs:82-92).
    __device__ float light_reciprocals(int c, int s, float x0, float x1, float ne) const {
        const int n = c + s + 1;
        float sum = 0.0f, next = 1.0f;
        for (int i = c; i < n; ++i) {
            const float l2c = i == n - 1 ? 1.0f : va(s - (i - c)).w;
            if (i == n - 1) sum += sdiv(next * ne, l2c);
            const float c2l = i == c ? x0 : (i == c + 1 ? x1 : vb(s + c + 2 - i).z);
            next *= sdiv(c2l, l2c);
            if (P.connections && i < n - 2) sum += next;
        }
        sum += next;
        return sum;
    }
};

// TraceLightPaths / TraceLightPath (BidirBase.Light.cs:197-219): SelectLight
// (:103-111) picks the background with BackgroundProbability, else an emitter;
// TraceEmitterPath / TraceBackgroundPath (:268-299) with RandomWalk.StartFromEmitter
// / StartFromBackground / ContinueWalk (RandomWalk.cs:82-128,150-217) and
// LightPathWalk (BidirBase.Light.cs:324-410) storing the vertices; with the light
// tracer on, every vertex is connected to the camera as it is stored
// (ConnectLightVertexToCamera, BidirBase.Light.cs:35-92, with
// ClassicBidir.LightTracerMis, ClassicBidir.cs:131-134): the connection draws
// no random numbers and reads only the vertices stored before it.
template <bool ENV>
__global__ void __launch_bounds__(kTraceBlock, 4) k_bidir_light(DevScene S, int scene, BidirParams P, float* image,
                                                             unsigned long long* stats) {
    __shared__ int stack_node[kStackNodeLds];
    __shared__ float stack_t[kStackLds * kTraceBlock];
    const SceneRec& R = S.recs[scene];
    const Env E = env_of(S, R);
    const f3 cam = mk3(R.cam.pos[0], R.cam.pos[1], R.cam.pos[2]);
    const float bg = ENV ? background_probability(R) : 0.0f;
    for (uint32_t idx = blockIdx.x * blockDim.x + threadIdx.x; idx < P.num_paths; idx += gridDim.x * blockDim.x) {
        BidirCtx<ENV> B{S, R, P, idx, bg};
        Rng rng{hash_seed(P.seed_light, idx, P.iteration)};
        const bool from_bg = bg > 0.0f && rng.next_float() <= bg;
        f3 prefix, approx = gray(1.0f);
        Surf prev{};  // the ancestor; for a background path first the ray origin, on no surface
        float pdf_dir, next_rev = 0.0f, ne = 0.0f;
        f3 o, d;
        float tmin;
        if (from_bg) {
            const f2 ppos = rng.next_float2();  // SampleBackground: primaryPos, then primaryDir
            const f2 pdir = rng.next_float2();
            EnvRay er = env_sample_ray(E, R, ppos, pdir);
            er.pdf *= bg;
            er.weight = er.weight / bg;
            if (er.pdf == 0.0f || is_zero(er.weight)) {
                P.lv_len[idx] = 0;
                continue;
            }
            P.lv_a[idx] = f4_of(er.o, 0.0f);
            P.lv_b[idx] = make_float4(0.0f, 0.0f, 0.0f, __int_as_float(-1));
            prefix = er.weight;
            prev.pos = er.o;
            pdf_dir = er.pdf;  // the first hit's pdf stays the ray's (per solid angle)
            o = er.o;
            d = er.d;
            tmin = 0.0f;
        } else {
            const DevEmitter e = S.emitters[R.emitter_off + (int)rng.next_int((uint32_t)R.num_emitters)];
            const float sel = (1.0f - bg) / (float)R.num_emitters;
            const f2 ppos = rng.next_float2();
            const f2 pdir = rng.next_float2();
            EmitterRay er = emitter_ray(S, e, ppos, pdir);
            er.pdf *= sel;
            ne = B.next_event_pdf(e);
            if (er.pdf == 0.0f || is_zero(er.weight)) {
                P.lv_len[idx] = 0;
                continue;
            }
            const f2 b0 = to_uniform_triangle(ppos);
            P.lv_a[idx] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            P.lv_b[idx] = make_float4(b0.x, b0.y, 0.0f, __int_as_float(e.tri));
            prefix = er.weight / sel;
            prev = er.p;
            pdf_dir = er.pdf;
            o = offset_point(er.p, er.dir);
            d = er.dir;
            tmin = er.p.err;
        }
        P.lv_ne[idx] = ne;
        int len = 1;
        // StartFromBackground traces and stores the first hit whatever MaxDepth, samples
        // its continuation without roulette, keeps its pdfs per solid angle and continues
        // with the BSDF weight alone (RandomWalk.cs:89-128); ContinueWalk then as for
        // emitter paths
        for (int depth = 1; depth < P.max_depth || (from_bg && depth == 1);) {
            const bool first_bg = from_bg && depth == 1;
            const HitOut h = traverse<false>(S, R.root, R.abs_max, o, d, tmin, INFINITY, stack_node + threadIdx.x,
                                             stack_t + threadIdx.x);
            wave_count(stats, kStatRays, true);
            wave_count(stats, kStatHits, h.li >= 0);
            if (h.li < 0) break;
            const Surf s = surface(S, h.li, h.u, h.v);
            const float pdf_from = first_bg ? pdf_dir : pdf_dir * area_to_solid_angle(prev.pos, s.pos, s.ng);
            if (pdf_from == 0.0f) break;
            const size_t at = (size_t)depth * P.num_paths + idx;
            P.lv_a[at] = f4_of(prefix, pdf_from);
            P.lv_b[at] = make_float4(h.u, h.v, next_rev, __int_as_float(h.li));
            len = depth + 1;
            if (first_bg) {  // NextEventPdf of the first edge: the map's pdf towards the origin
                ne = B.next_event_pdf_background(E, s.pos, prev.pos);
                P.lv_ne[idx] = ne;
            }
            uint32_t pix;
            float jac;
            if (P.mode == kModeClassic && P.light_tracer && depth + 1 >= P.min_depth &&
                camera_response(R.cam, s.pos, s.ng, &pix, &jac)) {
                const f3 p0 = offset_point(s, cam - s.pos);
                const bool occ = traverse<true>(S, R.root, R.abs_max, p0, cam - p0, 1e-5f, 1.0f - 1e-5f,
                                                stack_node + threadIdx.x, stack_t + threadIdx.x)
                                     .li >= 0;
                wave_count(stats, kStatShadow, true);
                wave_count(stats, kStatShadowTraced, true);
                wave_count(stats, kStatOccluded, occ);
                if (!occ) {
                    f3 to_cam = cam - s.pos;
                    to_cam = to_cam / length(to_cam);
                    const f3 to_anc = normalize(prev.pos - s.pos);
                    Shader sh;
                    make_shader(S, s, to_cam, &sh);  // the camera side: IsOnLightSubpath = false
                    float pdf_rev, unused;
                    f3 f = value_pdf_pair(sh, to_anc, &pdf_rev, &unused);
                    if (!is_zero(f)) {
                        f = f * sdiv(fabsf(dot(s.ns, to_anc)), fabsf(dot(s.ng, to_anc)));
                        if (!first_bg) pdf_rev *= area_to_solid_angle(s.pos, prev.pos, prev.ng);
                        // PdfEmit = invLensArea * jacobian, then the vertex's own reverse pdf
                        const float mis = srcp(
                            1.0f + sdiv(B.light_reciprocals(0, depth, 1.0f * jac, pdf_rev, ne), P.num_paths_f));
                        const f3 w = prefix * f * jac / P.num_paths_f;
                        const f3 v = mis * w;
                        const bool bad = !isfinite(average(v));
                        wave_count(stats, kStatNan, bad);
                        if (!bad) {
                            const f3 add = div_exact(v, P.n);
                            float* px = image + 3ull * pix;
                            atomicAdd(px, add.x);
                            atomicAdd(px + 1, add.y);
                            atomicAdd(px + 2, add.z);
                        }
                    }
                }
            }
            float q = 1.0f;
            if (!first_bg) {
                if (depth + 1 >= P.max_depth) break;
                q = depth > 4 ? clampf(average(approx), 0.05f, 0.95f) : 1.0f;
                if (rng.next_float() > q) break;
            }
            Shader sh;
            make_shader(S, s, -d, &sh);
            sh.light = true;
            const float pc = rng.next_float();
            const f2 pd = rng.next_float2();
            const BsdfSample bs = sample_bsdf<true>(sh, pc, pd);
            approx = approx * (bs.weight / q);
            // OnContinue
            next_rev = first_bg ? bs.pdf_rev : bs.pdf_rev * area_to_solid_angle(s.pos, prev.pos, prev.ng);
            if (first_bg && P.max_depth <= 1) break;
            if (bs.pdf == 0.0f || is_zero(bs.weight)) break;
            f3 w = bs.weight;
            if (!first_bg) {
                // RandomWalk.cs:193-204: the geometric instead of the shading cosine, and the
                // rendering equation's cosine ratio at this vertex of the light subpath
                w = w * sdiv(fabsf(dot(s.ng, bs.dir)), fabsf(dot(s.ns, bs.dir)));
                w = w * sdiv(fabsf(dot(s.ns, -d)), fabsf(dot(s.ng, -d)));
            }
            prefix = prefix * (w / q);
            ++depth;
            pdf_dir = bs.pdf;
            prev = s;
            o = offset_point(s, bs.dir);
            d = bs.dir;
            tmin = s.err;
        }
        P.lv_len[idx] = len;
    }
}

// ------------------------------------- ClassicBidir camera paths, wavefront -
// BidirBase.RenderPixel (BidirBase.Camera.cs:97-106) and the camera's RandomWalk
// (RandomWalk.cs:73-80,150-217, CameraRandomWalk BidirBase.Camera.cs:588-653) of
// MaxDepth + 1 vertices with ClassicBidir.OnCameraHit (ClassicBidir.cs:94-116):
// OnEmitterHit / OnBackgroundHit (BidirBase.Camera.cs:500-561), the connections
// to every vertex of the pixel's light path (BidirConnections / Connect,
// :192-307) and NumShadowRays next-event samples (PerformNextEventEstimation,
// :368-484), each weighted by the balance heuristic (ClassicBidir.cs:119-153).
// Per bounce k_bc_trace finds the live paths' closest hits, k_bc_shade shades
// them (the hit's emitter or background term, connections and next events
// written as shadow records that carry their weighted terms, the
// continuation), k_bc_shadow tests the records' segments and k_bc_gather adds
// each path's vertex value to its sum in the reference's order (emitter term,
// prefix times the visible connections' sum, each visible next event). The
// shadow rays run in a traversal kernel at the traversal's occupancy instead of
// inside the shading (camera pass 8.0 -> 5.35 ms per large-scene iteration),
// records whose term is zero are counted but not traced, and the live paths are
// compacted every bounce. The pixel value is added to the running mean after
// the light tracer's splats.
__device__ __forceinline__ uint32_t* bc_list(const BidirParams& P, int which) {
    return which ? P.bc_list1 : P.bc_list0;
}

template <bool ENV>
__global__ void __launch_bounds__(256) k_bc_init(DevScene S, int scene, BidirParams P) {
    const SceneRec& R = S.recs[scene];
    const DevCamera& C = R.cam;
    for (uint32_t pix = blockIdx.x * blockDim.x + threadIdx.x; pix < P.num_pixels; pix += gridDim.x * blockDim.x) {
        const uint32_t row = pix / (uint32_t)C.width, col = pix % (uint32_t)C.width;
        Rng rng{hash_seed(P.seed_camera, pix, P.iteration)};
        const f2 off = rng.next_float2();
        f3 o, d, raw;
        camera_ray(C, (float)col + off.x, (float)row + off.y, &o, &d, &raw);
        // CameraRaySample.Point (position, normal = Camera.Direction) and PdfRay
        P.bc_rng[pix] = rng.state;
        P.bc_a[pix] = make_float4(1.0f, 1.0f, 1.0f, pixel_jacobian(C, o + raw));
        P.bc_b[pix] = make_float4(1.0f, 1.0f, 1.0f, __int_as_float(1));
        P.bc_o[pix] = f4_of(o, 0.0f);
        P.bc_d[pix] = f4_of(d, 0.0f);
        P.bc_pp[pix] = f4_of(o, 0.0f);
        P.bc_pn[pix] = make_float4(-C.c2w[8], -C.c2w[9], -C.c2w[10], 0.0f);
        P.bc_L[pix] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        P.bc_list0[pix] = pix;
    }
}

__global__ void __launch_bounds__(kTraceBlock) __attribute__((amdgpu_waves_per_eu(SSPT_TRACE_WAVES)))
k_bc_trace(DevScene S, int scene, BidirParams P, int in, unsigned long long* stats) {
    __shared__ int stack_node[kStackNodeLds];
    __shared__ float stack_t[kStackLds * kTraceBlock];
    const SceneRec& R = S.recs[scene];
    const uint32_t n = P.bc_count[in];
    const uint32_t* list = bc_list(P, in);
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const uint32_t pix = list[i];
        const float4 o = P.bc_o[pix], d = P.bc_d[pix];
        const HitOut h = traverse<false>(S, R.root, R.abs_max, f3_of(o), f3_of(d), o.w, INFINITY,
                                         stack_node + threadIdx.x, stack_t + threadIdx.x);
        wave_count(stats, kStatRays, true);
        wave_count(stats, kStatHits, h.li >= 0);
        P.bc_hit[pix] = make_float4(h.u, h.v, h.t, __int_as_float(h.li));
    }
}

template <bool ENV>
__global__ void __launch_bounds__(kTraceBlock, 2) k_bc_shade(DevScene S, int scene, BidirParams P, int in,
                                                             unsigned long long* stats) {
    const SceneRec& R = S.recs[scene];
    const Env E = env_of(S, R);
    const float bg = ENV ? background_probability(R) : 0.0f;
    const uint32_t n = P.bc_count[in];
    const uint32_t* list = bc_list(P, in);
    uint32_t* next = bc_list(P, 1 - in);
    const int conn_slots = max(P.max_depth - 1, 0);
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const uint32_t pix = list[i];
        BidirCtx<ENV> B{S, R, P, pix, bg};
        Rng rng{P.bc_rng[pix]};
        const float4 a = P.bc_a[pix], b = P.bc_b[pix];
        f3 prefix = f3_of(a), approx = f3_of(b);
        const float pdf_dir = a.w;
        int depth = __float_as_int(b.w);
        const f3 d = f3_of(P.bc_d[pix]);
        const f3 prev_pos = f3_of(P.bc_pp[pix]), prev_n = f3_of(P.bc_pn[pix]);
        const float4 hit = P.bc_hit[pix];
        const int li = __float_as_int(hit.w);
        const size_t rec0 = (size_t)pix * P.bc_rec;
        f3 value = gray(0.0f);
        int nconn = 0, nnee = 0;
        bool conn_block = false;
        if (li < 0) {
            // CameraRandomWalk.OnInvalidHit / OnBackgroundHit with ClassicBidir.EmitterHitMis
            if (ENV) {
                const int c = depth;
                float mis = 1.0f;
                if (c > 1) {
                    const float x_prev = c - 2 >= 1 ? B.cam(c - 2).y : 0.0f;
                    float sum = 1.0f;
                    sum += sdiv(env_pdf(E, d) * bg, pdf_dir);
                    sum += sdiv(B.camera_reciprocals(c - 2, background_pdf(E, R, -d), x_prev), pdf_dir);
                    mis = srcp(sum);
                }
                value = prefix * (mis * env_radiance(E, d));
            }
            P.bc_v0[pix] = f4_of(value, __int_as_float(0));
            P.bc_v1[pix] = f4_of(prefix, __int_as_float(0));
            continue;
        }
        const Surf s = surface(S, li, hit.x, hit.y);
        const float pdf_from = pdf_dir * area_to_solid_angle(prev_pos, s.pos, s.ng);
        if (pdf_from == 0.0f) {
            P.bc_v0[pix] = f4_of(value, __int_as_float(0));
            P.bc_v1[pix] = f4_of(prefix, __int_as_float(0));
            continue;
        }
        const float rev_jac = area_to_solid_angle(s.pos, prev_pos, prev_n);
        Shader sh;
        make_shader(S, s, -d, &sh);
        const int c = depth;  // camera vertices, this one included
        P.cam[(size_t)(c - 1) * P.num_pixels + pix] = make_float2(pdf_from, 0.0f);
        const int ei = S.tri_emitter[li];
        if (ei >= 0 && depth >= P.min_depth) {
            const DevEmitter e = S.emitters[R.emitter_off + ei];
            const f3 emission = emitted(e, s.ns, -d);
            float mis = 1.0f;
            if (c > 1) {
                const float pdf_emit = B.emitter_pdf(e, s.ns, -d, rev_jac);
                const float x_prev = c - 2 >= 1 ? B.cam(c - 2).y : 0.0f;
                float sum = 1.0f;
                sum += sdiv(B.next_event_pdf(e), pdf_from);
                sum += sdiv(B.camera_reciprocals(c - 2, pdf_emit, x_prev), pdf_from);
                mis = srcp(sum);
            }
            value = value + prefix * (mis * emission);
        }
        const int light_len = P.lv_len[pix];
        if (depth < P.max_depth && P.connections && light_len > 1) {
            conn_block = true;
            const float ne = P.lv_ne[pix];
            const bool light_from_bg = from_background<ENV>(P, pix);
            Surf anc = B.vertex(0);
            for (int k = 1; k < light_len; ++k) {
                const Surf lv = B.vertex(k);
                const int total = k + c + 1;
                if (total <= P.max_depth && total >= P.min_depth) {
                    const size_t slot = rec0 + nconn++;
                    wave_count(stats, kStatShadow, true);
                    const f3 dd = s.pos - lv.pos;
                    const f3 to_anc = normalize(anc.pos - lv.pos);
                    const f3 dir = normalize(lv.pos - s.pos);  // camera vertex -> light vertex
                    Shader ls;
                    make_shader(S, lv, to_anc, &ls);
                    ls.light = true;
                    float pcl, pcr, plc, plr;
                    f3 f_light = value_pdf_pair(ls, -dir, &plc, &plr) * fabsf(dot(lv.ng, -dir));
                    f_light = f_light * sdiv(fabsf(dot(lv.ns, to_anc)), fabsf(dot(lv.ng, to_anc)));
                    const f3 f_cam = value_cos_pdf_pair(sh, dir, &pcl, &pcr);
                    pcr *= rev_jac;
                    pcl *= area_to_solid_angle(s.pos, lv.pos, lv.ng);
                    if (k > 1 || !light_from_bg) plr *= area_to_solid_angle(lv.pos, anc.pos, anc.ng);
                    plc *= area_to_solid_angle(lv.pos, s.pos, s.ng);
                    const float dist2 = dot(dd, dd);
                    f3 term = gray(0.0f);
                    if (!is_zero(f_cam) && !is_zero(f_light) && pcl != 0.0f && dist2 != 0.0f) {
                        float sum = 1.0f;
                        sum += B.camera_reciprocals(c - 1, plc, pcr);
                        sum += B.light_reciprocals(c, k, pcl, plr, ne);
                        const float mis = srcp(sum);
                        const f3 w = f3_of(B.va(k)) * f_light * f_cam / dist2 / 1.0f;
                        term = mis * w;
                    }
                    P.bc_rc[slot] = f4_of(term, 0.0f);
                    if (!is_zero(term)) {
                        const f3 p0 = offset_point(lv, dd), p1 = offset_point(s, -dd);
                        P.bc_r0[slot] = f4_of(p0, 0.0f);
                        P.bc_r1[slot] = f4_of(p1, 0.0f);
                        P.bc_shadow[atomicAdd(P.bc_count + 2, 1u)] = (uint32_t)slot;
                    }
                }
                anc = lv;
            }
        }
        if (depth < P.max_depth && depth + 1 >= P.min_depth) {
            for (int r = 0; r < P.num_shadow_rays; ++r) {
                const float u_bg = rng.next_float();  // drawn with or without a background
                if (ENV && u_bg < bg) {
                    EnvSample es = env_sample(E, rng.next_float2());
                    es.pdf *= bg;
                    es.weight = es.weight / bg;
                    if (es.pdf == 0.0f) continue;
                    const size_t slot = rec0 + conn_slots + nnee++;
                    wave_count(stats, kStatShadow, true);
                    float fwd, rev;
                    const f3 bc = value_cos_pdf_pair(sh, es.dir, &fwd, &rev);
                    rev *= rev_jac;
                    f3 term = gray(0.0f);
                    if (!(fwd == 0.0f || rev == 0.0f || is_zero(bc))) {
                        float sum = 1.0f;
                        sum += sdiv(fwd, es.pdf);
                        sum += sdiv(B.camera_reciprocals(c - 1, background_pdf(E, R, -es.dir), rev), es.pdf);
                        term = srcp(sum) * (es.weight * bc);
                    }
                    P.bc_rc[slot] = f4_of(term, 0.0f);
                    if (!is_zero(term)) {  // Raytracer.LeavesScene
                        P.bc_r0[slot] = f4_of(offset_point(s, es.dir), s.err);
                        P.bc_r1[slot] = f4_of(es.dir, 1.0f);
                        P.bc_shadow[atomicAdd(P.bc_count + 2, 1u)] = (uint32_t)slot;
                    }
                    continue;
                }
                if (R.num_emitters == 0) continue;
                const DevEmitter e = S.emitters[R.emitter_off + (int)rng.next_int((uint32_t)R.num_emitters)];
                const float prob = 1.0f / (float)R.num_emitters;
                const f2 bb = to_uniform_triangle(rng.next_float2());
                const Surf lp = surface(S, e.tri, bb.x, bb.y);
                float lpdf = e.inv_area;
                lpdf *= prob;
                lpdf *= (float)P.num_shadow_rays;
                lpdf *= (1.0f - bg);
                if (lpdf == 0.0f) continue;
                const size_t slot = rec0 + conn_slots + nnee++;
                wave_count(stats, kStatShadow, true);
                const f3 dd = lp.pos - s.pos;
                const f3 l2s = normalize(s.pos - lp.pos);
                const f3 emission = emitted(e, lp.ns, l2s);
                float fwd, rev;
                const f3 bc = value_cos_pdf_pair(sh, -l2s, &fwd, &rev);
                const float jac = area_to_solid_angle(s.pos, lp.pos, lp.ng);
                fwd *= jac;
                rev *= rev_jac;
                f3 term = gray(0.0f);
                if (!(is_zero(emission) || is_zero(bc) || jac == 0.0f || fwd == 0.0f || rev == 0.0f)) {
                    const float pdf_emit = B.emitter_pdf(e, lp.ns, l2s, area_to_solid_angle(lp.pos, s.pos, s.ng));
                    float sum = 1.0f;
                    sum += sdiv(fwd, lpdf);
                    sum += sdiv(B.camera_reciprocals(c - 1, pdf_emit, rev), lpdf);
                    term = srcp(sum) * (emission * bc * sdiv(jac, lpdf));
                }
                P.bc_rc[slot] = f4_of(term, 0.0f);
                if (!is_zero(term)) {
                    P.bc_r0[slot] = f4_of(offset_point(s, dd), 0.0f);
                    P.bc_r1[slot] = f4_of(offset_point(lp, -dd), 0.0f);
                    P.bc_shadow[atomicAdd(P.bc_count + 2, 1u)] = (uint32_t)slot;
                }
            }
        }
        P.bc_v0[pix] = f4_of(value, __int_as_float(nconn | (conn_block ? 0x10000 : 0)));
        P.bc_v1[pix] = f4_of(prefix, __int_as_float(nnee));
        // the continuation (RandomWalk.ContinueWalk)
        bool cont = depth + 1 < P.max_depth + 1;
        if (cont) {
            const float q = depth > 4 ? clampf(average(approx), 0.05f, 0.95f) : 1.0f;
            cont = !(rng.next_float() > q);
            if (cont) {
                const float pc = rng.next_float();
                const f2 pd = rng.next_float2();
                const BsdfSample bs = sample_bsdf<true>(sh, pc, pd);
                approx = approx * (bs.weight / q);
                P.cam[(size_t)(c - 1) * P.num_pixels + pix].y = bs.pdf_rev * rev_jac;  // OnContinue
                cont = !(bs.pdf == 0.0f || is_zero(bs.weight));
                if (cont) {
                    prefix = prefix * (bs.weight / q);
                    ++depth;
                    P.bc_a[pix] = f4_of(prefix, bs.pdf);
                    P.bc_b[pix] = f4_of(approx, __int_as_float(depth));
                    P.bc_o[pix] = f4_of(offset_point(s, bs.dir), s.err);
                    P.bc_d[pix] = f4_of(bs.dir, 0.0f);
                    P.bc_pp[pix] = f4_of(s.pos, 0.0f);
                    P.bc_pn[pix] = f4_of(s.ng, 0.0f);
                    P.bc_rng[pix] = rng.state;
                    next[atomicAdd(P.bc_count + (1 - in), 1u)] = pix;
                }
            }
        }
    }
}

__global__ void __launch_bounds__(kTraceBlock) __attribute__((amdgpu_waves_per_eu(SSPT_TRACE_WAVES)))
k_bc_shadow(DevScene S, int scene, BidirParams P, unsigned long long* stats) {
    __shared__ int stack_node[kStackNodeLds];
    __shared__ float stack_t[kStackLds * kTraceBlock];
    const SceneRec& R = S.recs[scene];
    const uint32_t n = P.bc_count[2];
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const uint32_t slot = P.bc_shadow[i];
        const float4 r0 = P.bc_r0[slot], r1 = P.bc_r1[slot];
        const f3 p0 = f3_of(r0), p1 = f3_of(r1);
        const bool to_infinity = r1.w != 0.0f;
        const bool occ = traverse<true>(S, R.root, R.abs_max, p0, to_infinity ? p1 : p1 - p0,
                                        to_infinity ? r0.w : 1e-5f, to_infinity ? INFINITY : 1.0f - 1e-5f,
                                        stack_node + threadIdx.x, stack_t + threadIdx.x)
                             .li >= 0;
        wave_count(stats, kStatShadowTraced, true);
        wave_count(stats, kStatOccluded, occ);
        P.bc_rc[slot].w = occ ? 0.0f : 1.0f;
    }
}

// The vertex's value in ClassicBidir.OnCameraHit's order: its emitter term,
// prefix times the sum of the visible connections, then each visible next event.
__global__ void __launch_bounds__(256) k_bc_gather(BidirParams P, int in) {
    const uint32_t n = P.bc_count[in];
    const uint32_t* list = bc_list(P, in);
    const int conn_slots = max(P.max_depth - 1, 0);
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const uint32_t pix = list[i];
        const float4 v0 = P.bc_v0[pix], v1 = P.bc_v1[pix];
        const int cw = __float_as_int(v0.w), nconn = cw & 0xFFFF, nnee = __float_as_int(v1.w);
        const f3 prefix = f3_of(v1);
        const size_t rec0 = (size_t)pix * P.bc_rec;
        f3 value = f3_of(v0);
        if (cw & 0x10000) {
            f3 conn = gray(0.0f);
            for (int r = 0; r < nconn; ++r) {
                const float4 rc = P.bc_rc[rec0 + r];
                if (rc.w != 0.0f) conn = conn + f3_of(rc);
            }
            value = value + prefix * conn;
        }
        for (int r = 0; r < nnee; ++r) {
            const float4 rc = P.bc_rc[rec0 + conn_slots + r];
            if (rc.w != 0.0f) value = value + prefix * f3_of(rc);
        }
        P.bc_L[pix] = f4_of(f3_of(P.bc_L[pix]) + value, 0.0f);
    }
}

__global__ void __launch_bounds__(256) k_bc_finish(BidirParams P, float* image, unsigned long long* stats) {
    for (uint32_t pix = blockIdx.x * blockDim.x + threadIdx.x; pix < P.num_pixels; pix += gridDim.x * blockDim.x) {
        const f3 L = f3_of(P.bc_L[pix]);
        const bool bad = !isfinite(average(L));
        wave_count(stats, kStatNan, bad);
        if (!bad) {
            const f3 add = div_exact(L, P.n);
            float* px = image + 3ull * pix;
            px[0] += add.x;
            px[1] += add.y;
            px[2] += add.z;
        }
    }
}

// ------------------------------- light vertex cache and merging (VCM) -------
// VertexCacheBidir (VertexCacheBidir.cs) and VertexConnectionAndMerging
// (VertexConnectionAndMerging.cs) over the same light-path cache: the light
// kernel is k_bidir_light (its light-tracer weight per mode); k_vcm_prepare
// numbers the cached vertices in path order (PathCache.cumPathLen) and builds
// the photon grid; k_vcm_camera follows the camera paths.

// The pdfs of one full path as BidirPathPdfs (BidirPathPdfs.cs) holds them,
// read where they live instead of gathered into arrays: C2L[j] (camera ->
// light) is camera vertex j's PdfFromAncestor for j < c, else light vertex
// (n + 1 - j)'s PdfReverseAncestor; L2C[j] (light -> camera) is 1 for
// j = n - 1, camera vertex j + 1's PdfToAncestor for j < c - 2, else light
// vertex (n - 1 - j)'s PdfFromAncestor; a technique overrides up to four
// entries with its own values (GatherCameraPdfs / GatherLightPdfs and the
// assignments after them, BidirBase.Camera.cs:192-270,368-524,
// BidirBase.Light.cs:35-92, VertexConnectionAndMerging.cs:329-425).
struct PdfView {
    const BidirParams* P;
    uint32_t pixel, lpath;
    int n, c;
    float ne;  // PdfNextEvent
    int no = 0;
    int oi[4];
    bool oc[4];  // the override is a C2L entry
    float ov[4];
    __device__ __forceinline__ void set_c2l(int j, float v) {
        oi[no] = j;
        oc[no] = true;
        ov[no++] = v;
    }
    __device__ __forceinline__ void set_l2c(int j, float v) {
        oi[no] = j;
        oc[no] = false;
        ov[no++] = v;
    }
    __device__ __forceinline__ float c2l(int j) const {
        for (int k = 0; k < no; ++k)
            if (oc[k] && oi[k] == j) return ov[k];
        if (j < c) return P->cam[(size_t)j * P->num_pixels + pixel].x;
        return P->lv_b[(size_t)(n + 1 - j) * P->num_paths + lpath].z;
    }
    __device__ __forceinline__ float l2c(int j) const {
        for (int k = 0; k < no; ++k)
            if (!oc[k] && oi[k] == j) return ov[k];
        if (j == n - 1) return 1.0f;
        if (j < c - 2) return P->cam[(size_t)(j + 1) * P->num_pixels + pixel].y;
        return P->lv_a[(size_t)(n - 1 - j) * P->num_paths + lpath].w;
    }
};

// VertexConnectionAndMerging.CorrelAwareRatios (VertexConnectionAndMerging.cs:580-646):
// the camera-side prefix products and light-side suffix products of one full
// path. Computed on the first lookup (the weights without correlation-aware
// merging never read them) into per-lane LDS (`lds`, stride kTraceBlock) for
// paths of up to kCorrelMax surface vertices, in the reference's order, so each
// lookup is two LDS reads; longer paths recompute the products per lookup.
constexpr int kCorrelMax = 16;
struct Correl {
    const PdfView* v;
    float accept_area;
    bool from_bg;  // the light end is the background: its last density stays per solid angle
    float* lds;    // cam[i] at lds[2 i * kTraceBlock], light[i] at lds[(2 i + 1) * kTraceBlock], or null
    mutable bool ready = false;
    __device__ __forceinline__ Correl(const PdfView& view, float dist_to_cam, bool from_background, float* lds_lane)
        : v(&view), from_bg(from_background), lds(lds_lane) {
        const float radius = dist_to_cam * 0.0174550649f;
        accept_area = radius * radius * kPi;
    }
    __device__ __forceinline__ float light_factor(int i, int nsv) const {
        float next = v->l2c(i) * accept_area;
        if (i == nsv - 1 && !from_bg) next *= accept_area;
        return fminf(next, 1.0f);
    }
    __device__ float operator[](int idx) const {
        if (idx == 0) return 1.0f;
        const int nsv = v->n - 1;
        float cam = 1.0f, light = 1.0f;
        if (lds && nsv <= kCorrelMax) {
            if (!ready) {
                float product = 1.0f;
                for (int i = 0; i < nsv; ++i) {
                    product *= fminf(v->c2l(i) * accept_area, 1.0f);
                    lds[2 * i * kTraceBlock] = product;
                }
                product = 1.0f;
                for (int i = nsv - 1; i >= 0; --i) {
                    product *= light_factor(i, nsv);
                    lds[(2 * i + 1) * kTraceBlock] = product;
                }
                ready = true;
            }
            cam = lds[2 * idx * kTraceBlock];
            light = lds[(2 * idx + 1) * kTraceBlock];
        } else {
            for (int i = 0; i <= idx; ++i) cam *= fminf(v->c2l(i) * accept_area, 1.0f);
            for (int i = nsv - 1; i >= idx; --i) light *= light_factor(i, nsv);
        }
        if (cam == 0.0f && light == 0.0f) return 1.0f;
        return sdiv(cam, cam + light - cam * light);
    }
};

// VertexCacheBidir.BidirSelectDensity (VertexCacheBidir.cs:71-82)
__device__ __forceinline__ float select_density(const BidirParams& P, int num_vertices) {
    if (num_vertices == 0) return 0.0f;
    return (1.0f / (float)num_vertices) * (float)(P.num_connections * (int)P.num_paths);
}

// CameraPathReciprocals: VertexCacheBidir.cs:210-223, VertexConnectionAndMerging.cs:839-889
__device__ float vcm_camera_reciprocals(const BidirParams& P, const PdfView& v, int last, float radius,
                                        const Correl& cr, float density) {
    const bool merging = P.mode == kModeVcm && P.merging;
    float sum = 0.0f, next = 1.0f;
    for (int i = last; i > 0; --i) {
        if (merging) {
            float accept = v.l2c(i) * kPi * radius * radius;
            if (P.correl) accept *= cr[i];
            sum += next * P.num_paths_f * accept;
        }
        next *= sdiv(v.l2c(i), v.c2l(i));
        if (P.num_connections > 0) sum += next * density;
    }
    if (P.light_tracer) sum += sdiv(next * v.l2c(0), v.c2l(0)) * P.num_paths_f;
    if (merging && P.merge_primary) sum += next * P.num_paths_f * v.l2c(0) * kPi * radius * radius;
    return sum;
}

// LightPathReciprocals: VertexCacheBidir.cs:225-237, VertexConnectionAndMerging.cs:891-929
__device__ float vcm_light_reciprocals(const BidirParams& P, const PdfView& v, int last, float radius,
                                       const Correl& cr, float density) {
    const bool merging = P.mode == kModeVcm && P.merging;
    const int n = v.n;
    float sum = 0.0f, next = 1.0f;
    for (int i = last + 1; i < n; ++i) {
        if (i == n - 1) sum += sdiv(next * v.ne, v.l2c(i));
        if (merging && i < n - 1 && (P.merge_primary || i > 0)) {
            float accept = v.c2l(i) * kPi * radius * radius;
            if (P.correl) accept *= cr[i];
            sum += next * P.num_paths_f * accept;
        }
        next *= sdiv(v.c2l(i), v.l2c(i));
        if (i < n - 2 && P.num_connections > 0) sum += next * density;
    }
    if (P.hitting) sum += next;
    return sum;
}

// The inclusive scan of n ints in three passes: per-block scans, one block
// scanning the block sums in chunks with a running carry, the carries added.
constexpr int kScanBlock = 1024;
uint32_t scan_tmp_words(uint32_t n) { return (n + kScanBlock - 1) / kScanBlock + 1; }
__device__ __forceinline__ int block_scan_incl(int v, int* lds) {
    lds[threadIdx.x] = v;
    __syncthreads();
    for (int off = 1; off < kScanBlock; off <<= 1) {
        const int add = threadIdx.x >= (unsigned)off ? lds[threadIdx.x - off] : 0;
        __syncthreads();
        lds[threadIdx.x] += add;
        __syncthreads();
    }
    const int r = lds[threadIdx.x];
    __syncthreads();
    return r;
}
__global__ void __launch_bounds__(kScanBlock) k_vscan_local(const int* in, int* out, uint32_t n, int* block_sums) {
    __shared__ int lds[kScanBlock];
    const uint32_t i = blockIdx.x * kScanBlock + threadIdx.x;
    const int r = block_scan_incl(i < n ? in[i] : 0, lds);
    if (i < n) out[i] = r;
    if (threadIdx.x == kScanBlock - 1) block_sums[blockIdx.x] = r;
}
__global__ void __launch_bounds__(kScanBlock) k_vscan_carry(int* block_sums, uint32_t nb) {
    __shared__ int lds[kScanBlock];
    __shared__ int total;
    int carry = 0;
    for (uint32_t base = 0; base < nb; base += kScanBlock) {
        const uint32_t i = base + threadIdx.x;
        const int v = i < nb ? block_sums[i] : 0;
        const int r = block_scan_incl(v, lds);
        if (i < nb) block_sums[i] = carry + r - v;  // exclusive: the sum of the blocks before block i
        if (threadIdx.x == kScanBlock - 1) total = r;
        __syncthreads();
        carry += total;
        __syncthreads();
    }
}
__global__ void __launch_bounds__(kScanBlock) k_vscan_add(int* out, uint32_t n, const int* block_sums) {
    const uint32_t i = blockIdx.x * kScanBlock + threadIdx.x;
    if (i < n) out[i] += block_sums[blockIdx.x];
}

// The photon map (VertexConnectionAndMerging.ProcessPathCache, :297-316): the
// cached vertices of depth >= 1 with non-zero weight, hashed by grid cell.
__device__ __forceinline__ int3 grid_cell(f3 p, float h) {
    return make_int3((int)floorf(p.x / h), (int)floorf(p.y / h), (int)floorf(p.z / h));
}
__device__ __forceinline__ uint32_t cell_bucket(int3 c, uint32_t mask) {
    return ((uint32_t)c.x * 73856093u ^ (uint32_t)c.y * 19349663u ^ (uint32_t)c.z * 83492791u) & mask;
}
template <bool SCATTER>
__global__ void __launch_bounds__(256) k_photon_grid(DevScene S, BidirParams P) {
    const uint64_t total = (uint64_t)(P.max_depth - 1) * P.num_paths;
    for (uint64_t t = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; t < total;
         t += (uint64_t)gridDim.x * blockDim.x) {
        const int k = (int)(t / P.num_paths) + 1;
        const uint32_t p = (uint32_t)(t % P.num_paths);
        if (k >= P.lv_len[p]) continue;
        const size_t at = (size_t)k * P.num_paths + p;
        const float4 a = P.lv_a[at];
        if (a.x == 0.0f && a.y == 0.0f && a.z == 0.0f) continue;
        const float4 b = P.lv_b[at];
        const f3 pos = surface(S, __float_as_int(b.w), b.x, b.y).pos;
        const uint32_t bucket = cell_bucket(grid_cell(pos, P.cell), P.table_mask);
        if (!SCATTER) {
            atomicAdd(P.bucket_fill + bucket, 1u);
        } else {
            const uint32_t slot = atomicAdd(P.bucket_fill + bucket, 1u);
            const int g = (p == 0 ? 0 : P.cum[p - 1]) + k;  // global vertex index, the photon's order
            P.photons[P.bucket_start[bucket] + slot] = f4_of(pos, __int_as_float(g));
        }
    }
}
__global__ void k_fill_u32(uint32_t* p, uint64_t n, uint32_t v) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = v;
}
// The light vertex of global index g (PathCache.GetVertex, PathCache.cs:40-56):
// its path (the first with cum > g) and its index along the path.
__device__ __forceinline__ void global_vertex(const BidirParams& P, int g, uint32_t* path, int* k) {
    uint32_t lo = 0, len = P.num_paths;
    while (len > 0) {
        const uint32_t half = len >> 1;
        if (P.cum[lo + half] <= g) {
            lo += half + 1;
            len -= half + 1;
        } else {
            len = half;
        }
    }
    *path = lo;
    *k = g - (lo == 0 ? 0 : P.cum[lo - 1]);
}

// NearestNeighborSearch.ForAllNearest (TinyEmbree 1.2.0) as restated: the
// (at most) kmax photons with the smallest squared distances (fp32,
// dx*dx + dy*dy + dz*dz) below r2, ties to the smaller global index, in
// ascending order. Returns the count.
__device__ __forceinline__ void consider_photon(float4 ph, f3 q, float r2, int kmax, float* d2s, int* ids,
                                                int* found) {
    const float dx = ph.x - q.x, dy = ph.y - q.y, dz = ph.z - q.z;
    const float d2 = dx * dx + dy * dy + dz * dz;
    if (!(d2 < r2)) return;
    const int id = __float_as_int(ph.w);
    if (*found == kmax && (d2 > d2s[kmax - 1] || (d2 == d2s[kmax - 1] && id > ids[kmax - 1]))) return;
    int at = *found < kmax ? (*found)++ : kmax - 1;
    while (at > 0 && (d2s[at - 1] > d2 || (d2s[at - 1] == d2 && ids[at - 1] > id))) {
        d2s[at] = d2s[at - 1];
        ids[at] = ids[at - 1];
        --at;
    }
    d2s[at] = d2;
    ids[at] = id;
}
// The query visits the cells of its box in shells around q's own cell, nearest
// first: shell s holds the cells whose index differs from q's by s on some
// axis. A photon of shell s lies more than (s - 3) * h from q (one cell for q's
// place in its cell, one for each rounded cell index), so once k photons are
// found closer than that no later shell can change the result, ties included;
// the search stops there instead of visiting the whole box (a grazing primary
// hit has a footprint of many cells). A box still unfinished after
// kMaxShellCells cells, or wider than 4096 cells, scans every photon instead.
constexpr int kMaxShellCells = 32768;
__device__ int nearest_photons(const BidirParams& P, f3 q, float radius, float r2, int kmax, float* d2s, int* ids) {
    int found = 0;
    const float h = P.cell;
    auto every_photon = [&]() {
        found = 0;
        const uint32_t end = P.bucket_start[P.table_mask + 1];
        for (uint32_t e = 0; e < end; ++e) consider_photon(P.photons[e], q, r2, kmax, d2s, ids, &found);
        return found;
    };
    const float span = 2.0f * radius / h + 2.0f;  // cells per axis the query box can touch
    if (!(span <= 4096.0f)) return every_photon();  // also a non-finite radius
    const int3 lo = grid_cell(mk3(q.x - radius, q.y - radius, q.z - radius), h);
    const int3 hi = grid_cell(mk3(q.x + radius, q.y + radius, q.z + radius), h);
    const int3 c0 = grid_cell(q, h);
    const int smax = max(max(max(c0.x - lo.x, hi.x - c0.x), max(c0.y - lo.y, hi.y - c0.y)),
                         max(c0.z - lo.z, hi.z - c0.z));
    auto visit = [&](int x, int y, int z) {
        const uint32_t b = cell_bucket(make_int3(x, y, z), P.table_mask);
        const uint32_t end = P.bucket_start[b + 1];
        for (uint32_t e = P.bucket_start[b]; e < end; ++e) {
            const float4 ph = P.photons[e];
            const int3 pc = grid_cell(f3_of(ph), h);
            if (pc.x != x || pc.y != y || pc.z != z) continue;  // another cell in the same bucket
            consider_photon(ph, q, r2, kmax, d2s, ids, &found);
        }
    };
    int visited = 0;
    for (int sh = 0; sh <= smax; ++sh) {
        if (found == kmax && sh >= 4) {
            const float bound = (float)(sh - 3) * h;
            if (d2s[kmax - 1] < bound * bound * 0.99999f) break;
        }
        if (visited > kMaxShellCells) return every_photon();
        const int x0 = max(lo.x, c0.x - sh), x1 = min(hi.x, c0.x + sh);
        const int y0 = max(lo.y, c0.y - sh), y1 = min(hi.y, c0.y + sh);
        const int z0 = max(lo.z, c0.z - sh), z1 = min(hi.z, c0.z + sh);
        for (int x = x0; x <= x1; ++x)
            for (int y = y0; y <= y1; ++y) {
                if (abs(x - c0.x) == sh || abs(y - c0.y) == sh) {  // a face of the shell: its whole z range
                    for (int z = z0; z <= z1; ++z) visit(x, y, z);
                    visited += z1 - z0 + 1;
                } else {  // inside in x and y: the shell's two z caps
                    if (c0.z - sh >= lo.z) visit(x, y, c0.z - sh), ++visited;
                    if (c0.z + sh <= hi.z) visit(x, y, c0.z + sh), ++visited;
                }
            }
    }
    return found;
}

// SplatLightVertices (BidirBase.Light.cs:29-92) of VertexCacheBidir / VCM, after
// the cache is numbered: the light tracer's weight counts the connections,
// whose density needs the vertex count (LightTracerMis, VertexCacheBidir.cs:183-186,
// VertexConnectionAndMerging.cs:735-766).
template <bool ENV>
__global__ void __launch_bounds__(kTraceBlock) k_vcm_splat(DevScene S, int scene, BidirParams P, float* image,
                                                           unsigned long long* stats) {
    __shared__ int stack_node[kStackNodeLds];
    __shared__ float stack_t[kStackLds * kTraceBlock];
    __shared__ float correl_lds[2 * kCorrelMax * kTraceBlock];
    const SceneRec& R = S.recs[scene];
    const f3 cam = mk3(R.cam.pos[0], R.cam.pos[1], R.cam.pos[2]);
    const int num_vertices = P.num_paths > 0 ? P.cum[P.num_paths - 1] : 0;
    const float density = select_density(P, num_vertices);
    for (uint32_t idx = blockIdx.x * blockDim.x + threadIdx.x; idx < P.num_paths; idx += gridDim.x * blockDim.x) {
        const int len = P.lv_len[idx];
        if (len < 2) continue;
        const float ne = P.lv_ne[idx];
        const bool from_bg = from_background<ENV>(P, idx);
        Surf prev = light_vertex<ENV>(S, P, 0, idx);
        for (int depth = 1; depth < len; ++depth) {
            const size_t at = (size_t)depth * P.num_paths + idx;
            const float4 a = P.lv_a[at];
            const float4 b = P.lv_b[at];
            const Surf s = surface(S, __float_as_int(b.w), b.x, b.y);
            uint32_t pix;
            float jac;
            if (depth + 1 >= P.min_depth && camera_response(R.cam, s.pos, s.ng, &pix, &jac)) {
                const f3 p0 = offset_point(s, cam - s.pos);
                const bool occ = traverse<true>(S, R.root, R.abs_max, p0, cam - p0, 1e-5f, 1.0f - 1e-5f,
                                                stack_node + threadIdx.x, stack_t + threadIdx.x)
                                     .li >= 0;
                wave_count(stats, kStatShadow, true);
                wave_count(stats, kStatShadowTraced, true);
                wave_count(stats, kStatOccluded, occ);
                if (!occ) {
                    f3 to_cam = cam - s.pos;
                    const float dist = length(to_cam);
                    to_cam = to_cam / dist;
                    const f3 to_anc = normalize(prev.pos - s.pos);
                    Shader sh;
                    make_shader(S, s, to_cam, &sh);
                    float pdf_rev, unused;
                    f3 f = value_pdf_pair(sh, to_anc, &pdf_rev, &unused);
                    if (!is_zero(f)) {
                        f = f * sdiv(fabsf(dot(s.ns, to_anc)), fabsf(dot(s.ng, to_anc)));
                        if (depth > 1 || !from_bg) pdf_rev *= area_to_solid_angle(s.pos, prev.pos, prev.ng);
                        PdfView v{&P, 0u, idx, depth + 1, 0, ne};
                        v.set_c2l(0, 1.0f * jac);
                        v.set_c2l(1, pdf_rev);
                        float mis;
                        if (P.mode == kModeVcm) {
                            const Correl cr(v, dist, from_bg, correl_lds + threadIdx.x);
                            const float footprint = sqrtf(1.0f / v.c2l(0));
                            float sum = vcm_light_reciprocals(P, v, -1, footprint, cr, density);
                            sum = sdiv(sum, P.num_paths_f);
                            sum += 1.0f;
                            mis = srcp(sum);
                        } else {
                            const Correl cr(v, dist, from_bg, correl_lds + threadIdx.x);
                            mis = srcp(1.0f + sdiv(vcm_light_reciprocals(P, v, -1, 0.0f, cr, density),
                                                   P.num_paths_f));
                        }
                        const f3 w = f3_of(a) * f * jac / P.num_paths_f;
                        const f3 val = mis * w;
                        const bool bad = !isfinite(average(val));
                        wave_count(stats, kStatNan, bad);
                        if (!bad) {
                            const f3 add = div_exact(val, P.n);
                            float* px = image + 3ull * pix;
                            atomicAdd(px, add.x);
                            atomicAdd(px + 1, add.y);
                            atomicAdd(px + 2, add.z);
                        }
                    }
                }
            }
            prev = s;
        }
    }
}

// BidirBase.RenderPixel (BidirBase.Camera.cs:97-106) and the camera's RandomWalk
// with OnCameraHit of VertexCacheBidir (VertexCacheBidir.cs:143-168) /
// VertexConnectionAndMerging (VertexConnectionAndMerging.cs:525-572): emitter
// hits, NumConnections connections to vertices drawn from the whole cache,
// merging with the nearest photons (VCM) and NumShadowRays next-event samples,
// each weighted against all enabled techniques (VertexCacheBidir.cs:171-237,
// VertexConnectionAndMerging.cs:648-929).
template <bool ENV>
__global__ void __launch_bounds__(kTraceBlock, 3) k_vcm_camera(DevScene S, int scene, BidirParams P, float* image,
                                                            unsigned long long* stats) {
    __shared__ int stack_node[kStackNodeLds];
    __shared__ float stack_t[kStackLds * kTraceBlock];
    __shared__ float correl_lds[2 * kCorrelMax * kTraceBlock];
    const SceneRec& R = S.recs[scene];
    const DevCamera& C = R.cam;
    const bool vcm = P.mode == kModeVcm;
    const int num_vertices = P.num_paths > 0 ? P.cum[P.num_paths - 1] : 0;
    const float density = select_density(P, num_vertices);
    const Env E = env_of(S, R);
    const float bg = ENV ? background_probability(R) : 0.0f;
    const float pmf_light = (1.0f - bg) / (float)R.num_emitters;  // unused without emitters
    for (uint32_t pix = blockIdx.x * blockDim.x + threadIdx.x; pix < P.num_pixels; pix += gridDim.x * blockDim.x) {
        const uint32_t row = pix / (uint32_t)C.width, col = pix % (uint32_t)C.width;
        Rng rng{hash_seed(P.seed_camera, pix, P.iteration)};
        const f2 off = rng.next_float2();
        f3 o, d, raw;
        camera_ray(C, (float)col + off.x, (float)row + off.y, &o, &d, &raw);
        float tmin = 0.0f;
        f3 prev_pos = o;
        f3 prev_n = mk3(-C.c2w[8], -C.c2w[9], -C.c2w[10]);
        float pdf_dir = pixel_jacobian(C, o + raw);
        float footprint = 0.0f, dist0 = 0.0f;
        f3 L = gray(0.0f), prefix = gray(1.0f), approx = gray(1.0f);
        for (int depth = 1; depth < P.max_depth + 1;) {
            const HitOut h = traverse<false>(S, R.root, R.abs_max, o, d, tmin, INFINITY, stack_node + threadIdx.x,
                                             stack_t + threadIdx.x);
            wave_count(stats, kStatRays, true);
            wave_count(stats, kStatHits, h.li >= 0);
            if (h.li < 0) {
                // OnInvalidHit / OnBackgroundHit (BidirBase.Camera.cs:532-561,604-612) with
                // EmitterHitMis (VertexCacheBidir.cs:171-180, VertexConnectionAndMerging.cs:700-732),
                // whatever EnableHitting and MinDepth say
                if (ENV) {
                    const int c = depth;
                    P.cam[(size_t)(c - 1) * P.num_pixels + pix] = make_float2(pdf_dir, 0.0f);
                    float mis = 1.0f;
                    if (c > 1) {
                        PdfView v{&P, pix, 0u, c, c, env_pdf(E, d) * bg};
                        v.set_l2c(c - 2, background_pdf(E, R, -d));
                        const Correl cr(v, dist0, true, correl_lds + threadIdx.x);
                        float sum = 1.0f;
                        sum += sdiv(v.ne, pdf_dir);
                        sum += sdiv(vcm_camera_reciprocals(P, v, c - 2, footprint, cr, density), pdf_dir);
                        mis = srcp(sum);
                    }
                    L = L + prefix * (mis * env_radiance(E, d));
                }
                break;
            }
            const Surf s = surface(S, h.li, h.u, h.v);
            const float pdf_from = pdf_dir * area_to_solid_angle(prev_pos, s.pos, s.ng);
            if (pdf_from == 0.0f) break;
            const float rev_jac = area_to_solid_angle(s.pos, prev_pos, prev_n);
            Shader sh;
            make_shader(S, s, -d, &sh);
            if (depth == 1) {
                footprint = sqrtf(1.0f / pdf_from);  // CameraRandomWalk.OnHit: the pixel footprint
                dist0 = h.t;
            }
            const int c = depth;
            P.cam[(size_t)(c - 1) * P.num_pixels + pix] = make_float2(pdf_from, 0.0f);
            f3 value = gray(0.0f);
            // emitter hits
            const int ei = S.tri_emitter[h.li];
            const bool hitting = vcm ? (P.hitting || depth == 1) : P.hitting != 0;
            if (ei >= 0 && hitting && depth >= P.min_depth) {
                const DevEmitter e = S.emitters[R.emitter_off + ei];
                const f3 emission = emitted(e, s.ns, -d);
                float mis = 1.0f;
                if (c > 1) {
                    PdfView v{&P, pix, 0u, c, c, e.inv_area * (1.0f / (float)R.num_emitters) * (1.0f - bg) *
                                                     (float)P.num_shadow_rays};
                    v.set_l2c(c - 2, emitter_pdf_ray(e, s.ns, -d) * rev_jac * pmf_light);
                    const Correl cr(v, dist0, false, correl_lds + threadIdx.x);
                    const float pdf_this = pdf_from;
                    float sum = 1.0f;
                    sum += sdiv(v.ne, pdf_this);
                    sum += sdiv(vcm_camera_reciprocals(P, v, c - 2, footprint, cr, density), pdf_this);
                    mis = srcp(sum);
                }
                value = value + prefix * (mis * emission);
            }
            if (depth < P.max_depth) {
                // connections (BidirConnections with VertexSelector, BidirBase.Camera.cs:279-295)
                for (int r = 0; r < P.num_connections; ++r) {
                    const int g = (int)rng.next_int((uint32_t)num_vertices);
                    if (g <= 0) continue;
                    uint32_t lp;
                    int k;
                    global_vertex(P, g, &lp, &k);
                    if (k < 1) continue;  // a light point
                    const int total = k + c + 1;
                    if (total > P.max_depth || total < P.min_depth) continue;
                    const float4 la = P.lv_a[(size_t)k * P.num_paths + lp];
                    const float4 lb = P.lv_b[(size_t)k * P.num_paths + lp];
                    const Surf lv = surface(S, __float_as_int(lb.w), lb.x, lb.y);
                    const Surf anc = light_vertex<ENV>(S, P, k - 1, lp);
                    const bool lp_bg = from_background<ENV>(P, lp);
                    const f3 dd = s.pos - lv.pos;
                    const f3 p0 = offset_point(lv, dd), p1 = offset_point(s, -dd);
                    const bool occ = traverse<true>(S, R.root, R.abs_max, p0, p1 - p0, 1e-5f, 1.0f - 1e-5f,
                                                    stack_node + threadIdx.x, stack_t + threadIdx.x)
                                         .li >= 0;
                    wave_count(stats, kStatShadow, true);
                    wave_count(stats, kStatShadowTraced, true);
                    wave_count(stats, kStatOccluded, occ);
                    if (occ) continue;
                    const f3 to_anc = normalize(anc.pos - lv.pos);
                    const f3 dir = normalize(lv.pos - s.pos);
                    Shader ls;
                    make_shader(S, lv, to_anc, &ls);
                    ls.light = true;
                    float pcl, pcr, plc, plr;
                    f3 f_light = value_pdf_pair(ls, -dir, &plc, &plr) * fabsf(dot(lv.ng, -dir));
                    f_light = f_light * sdiv(fabsf(dot(lv.ns, to_anc)), fabsf(dot(lv.ng, to_anc)));
                    const f3 f_cam = value_cos_pdf_pair(sh, dir, &pcl, &pcr);
                    pcr *= rev_jac;
                    pcl *= area_to_solid_angle(s.pos, lv.pos, lv.ng);
                    if (k > 1 || !lp_bg) plr *= area_to_solid_angle(lv.pos, anc.pos, anc.ng);
                    plc *= area_to_solid_angle(lv.pos, s.pos, s.ng);
                    const float dist2 = dot(dd, dd);
                    if (is_zero(f_cam) || is_zero(f_light) || pcl == 0.0f || dist2 == 0.0f) continue;
                    PdfView v{&P, pix, lp, c + k + 1, c, P.lv_ne[lp]};
                    if (c > 1) v.set_l2c(c - 2, pcr);
                    v.set_l2c(c - 1, plc);
                    v.set_c2l(c, pcl);
                    v.set_c2l(c + 1, plr);
                    const Correl cr(v, dist0, lp_bg, correl_lds + threadIdx.x);
                    float sum = 1.0f;
                    sum += sdiv(vcm_camera_reciprocals(P, v, c - 1, footprint, cr, density), density);
                    sum += sdiv(vcm_light_reciprocals(P, v, c - 1, footprint, cr, density), density);
                    const float mis = srcp(sum);
                    const f3 w = f3_of(la) * f_light * f_cam / dist2 / density;
                    value = value + prefix * (mis * w);
                }
                // merging (PerformMerging / Merge, VertexConnectionAndMerging.cs:329-516)
                if (vcm && P.merging && (P.merge_primary || c > 1)) {
                    const float r2 = footprint * footprint;
                    float d2s[kMaxPhotons];
                    int ids[kMaxPhotons];
                    const int found = nearest_photons(P, s.pos, footprint, r2, P.max_photons, d2s, ids);
                    f3 est = gray(0.0f);
                    const float furthest = found > 0 ? sqrtf(d2s[found - 1]) : 0.0f;
                    for (int f = 0; f < found; ++f) {
                        const float distance = sqrtf(d2s[f]);
                        const float dist_sqr = distance * distance;
                        const float radius_sqr = found == P.max_photons ? furthest * furthest : r2;
                        uint32_t lp;
                        int k;
                        global_vertex(P, ids[f], &lp, &k);
                        const int total = c + k;
                        if (total > P.max_depth || total < P.min_depth) continue;
                        const float4 la = P.lv_a[(size_t)k * P.num_paths + lp];
                        const float4 lb = P.lv_b[(size_t)k * P.num_paths + lp];
                        const Surf ph = surface(S, __float_as_int(lb.w), lb.x, lb.y);
                        if (fabsf(dot(s.ng, ph.ng)) < 0.4f) continue;
                        const Surf anc = light_vertex<ENV>(S, P, k - 1, lp);
                        const bool lp_bg = from_background<ENV>(P, lp);
                        const bool first_from_bg = k == 1 && lp_bg;
                        const f3 to_anc = normalize(anc.pos - s.pos);
                        float pdf_light_rev, pdf_cam_rev;
                        f3 bsdf = value_pdf_pair(sh, to_anc, &pdf_light_rev, &pdf_cam_rev);
                        bsdf = bsdf * sdiv(fabsf(dot(s.ns, to_anc)), fabsf(dot(ph.ng, to_anc)));
                        const f3 contrib = f3_of(la) * bsdf / P.num_paths_f;
                        if (is_zero(contrib)) continue;
                        if (fabsf(dot(to_anc, s.ng)) < 1e-4f) continue;
                        pdf_cam_rev *= rev_jac;
                        // at the first hit from the background the pdf stays per solid angle
                        if (!first_from_bg) pdf_light_rev *= area_to_solid_angle(s.pos, anc.pos, anc.ng);
                        if (pdf_cam_rev == 0.0f || pdf_light_rev == 0.0f) continue;
                        // NextEventPdf(shader point, ancestor) for a photon of depth 1
                        const float ne_m = first_from_bg ? env_pdf(E, anc.pos - s.pos) * bg * (float)P.num_shadow_rays
                                                         : P.lv_ne[lp];
                        PdfView v{&P, pix, lp, c + k, c, ne_m};
                        if (c > 1) v.set_l2c(c - 2, pdf_cam_rev);
                        v.set_c2l(c, pdf_light_rev);
                        const Correl cr(v, dist0, lp_bg, correl_lds + threadIdx.x);
                        // MergeMis (:648-697)
                        float approx_m = v.l2c(c - 1) * kPi * footprint * footprint * P.num_paths_f;
                        if (P.correl) approx_m *= cr[c - 1];
                        float mis = 0.0f;
                        if (approx_m != 0.0f) {
                            float sum = 0.0f;
                            sum += sdiv(vcm_camera_reciprocals(P, v, c - 1, footprint, cr, density), approx_m);
                            sum += sdiv(vcm_light_reciprocals(P, v, c - 1, footprint, cr, density), approx_m);
                            if (k > 1 && P.num_connections > 0) sum += sdiv(density, approx_m);
                            mis = srcp(sum);
                        }
                        const float kernel = sdiv(2.0f * (radius_sqr - dist_sqr), kPi * radius_sqr * radius_sqr);
                        est = est + contrib * kernel * mis;
                    }
                    value = value + prefix * est;
                }
            }
            if (depth < P.max_depth && depth + 1 >= P.min_depth) {
                for (int r = 0; r < P.num_shadow_rays; ++r) {
                    const float u_bg = rng.next_float();  // drawn with or without a background
                    if (ENV && u_bg < bg) {
                        // the background branch (BidirBase.Camera.cs:381-418), NextEventMis with
                        // isBackground (VertexCacheBidir.cs:200-207, VertexConnectionAndMerging.cs:802-836)
                        EnvSample es = env_sample(E, rng.next_float2());
                        es.pdf *= bg;
                        es.weight = es.weight / bg;
                        if (es.pdf == 0.0f) continue;
                        const bool occ = traverse<true>(S, R.root, R.abs_max, offset_point(s, es.dir), es.dir, s.err,
                                                        INFINITY, stack_node + threadIdx.x, stack_t + threadIdx.x)
                                             .li >= 0;  // Raytracer.LeavesScene
                        wave_count(stats, kStatShadow, true);
                        wave_count(stats, kStatShadowTraced, true);
                        wave_count(stats, kStatOccluded, occ);
                        if (occ) continue;
                        float fwd, rev;
                        const f3 bc = value_cos_pdf_pair(sh, es.dir, &fwd, &rev);
                        rev *= rev_jac;
                        if (fwd == 0.0f || rev == 0.0f || is_zero(bc)) continue;
                        PdfView v{&P, pix, 0u, c + 1, c, es.pdf};
                        if (c > 1) v.set_l2c(c - 2, rev);
                        v.set_l2c(c - 1, background_pdf(E, R, -es.dir));
                        v.set_c2l(c, fwd);
                        const Correl cr(v, dist0, true, correl_lds + threadIdx.x);
                        float sum = 1.0f;
                        if (P.hitting) sum += sdiv(fwd, es.pdf);
                        sum += sdiv(vcm_camera_reciprocals(P, v, c - 1, footprint, cr, density), es.pdf);
                        value = value + prefix * (srcp(sum) * (es.weight * bc));
                        continue;
                    }
                    if (R.num_emitters == 0) continue;
                    const DevEmitter e = S.emitters[R.emitter_off + (int)rng.next_int((uint32_t)R.num_emitters)];
                    const float prob = 1.0f / (float)R.num_emitters;
                    const f2 b = to_uniform_triangle(rng.next_float2());
                    const Surf lp = surface(S, e.tri, b.x, b.y);
                    float lpdf = e.inv_area;
                    lpdf *= prob;
                    lpdf *= (float)P.num_shadow_rays;
                    lpdf *= (1.0f - bg);
                    if (lpdf == 0.0f) continue;
                    const f3 dd = lp.pos - s.pos;
                    const f3 p0 = offset_point(s, dd), p1 = offset_point(lp, -dd);
                    const bool occ = traverse<true>(S, R.root, R.abs_max, p0, p1 - p0, 1e-5f, 1.0f - 1e-5f,
                                                    stack_node + threadIdx.x, stack_t + threadIdx.x)
                                         .li >= 0;
                    wave_count(stats, kStatShadow, true);
                    wave_count(stats, kStatShadowTraced, true);
                    wave_count(stats, kStatOccluded, occ);
                    if (occ) continue;
                    const f3 l2s = normalize(s.pos - lp.pos);
                    const f3 emission = emitted(e, lp.ns, l2s);
                    float fwd, rev;
                    const f3 bc = value_cos_pdf_pair(sh, -l2s, &fwd, &rev);
                    const float jac = area_to_solid_angle(s.pos, lp.pos, lp.ng);
                    fwd *= jac;
                    rev *= rev_jac;
                    if (is_zero(emission) || is_zero(bc) || jac == 0.0f || fwd == 0.0f || rev == 0.0f) continue;
                    PdfView v{&P, pix, 0u, c + 1, c, lpdf};
                    if (c > 1) v.set_l2c(c - 2, rev);
                    v.set_l2c(c - 1, emitter_pdf_ray(e, lp.ns, l2s) * area_to_solid_angle(lp.pos, s.pos, s.ng) *
                                         pmf_light);
                    v.set_c2l(c, fwd);
                    const Correl cr(v, dist0, false, correl_lds + threadIdx.x);
                    float sum = 1.0f;
                    if (P.hitting) sum += sdiv(fwd, lpdf);
                    sum += sdiv(vcm_camera_reciprocals(P, v, c - 1, footprint, cr, density), lpdf);
                    const float mis = srcp(sum);
                    const f3 w = emission * bc * sdiv(jac, lpdf);
                    value = value + prefix * (mis * w);
                }
            }
            L = L + value;
            if (depth + 1 >= P.max_depth + 1) break;
            const float q = depth > 4 ? clampf(average(approx), 0.05f, 0.95f) : 1.0f;
            if (rng.next_float() > q) break;
            const float pc = rng.next_float();
            const f2 pd = rng.next_float2();
            const BsdfSample bs = sample_bsdf<true>(sh, pc, pd);
            approx = approx * (bs.weight / q);
            P.cam[(size_t)(c - 1) * P.num_pixels + pix].y = bs.pdf_rev * rev_jac;  // OnContinue
            if (bs.pdf == 0.0f || is_zero(bs.weight)) break;
            prefix = prefix * (bs.weight / q);
            ++depth;
            pdf_dir = bs.pdf;
            prev_pos = s.pos;
            prev_n = s.ng;
            o = offset_point(s, bs.dir);
            d = bs.dir;
            tmin = s.err;
        }
        const bool bad = !isfinite(average(L));
        wave_count(stats, kStatNan, bad);
        if (!bad) {
            const f3 add = div_exact(L, P.n);
            float* px = image + 3ull * pix;
            px[0] += add.x;
            px[1] += add.y;
            px[2] += add.z;
        }
    }
}


__global__ void __launch_bounds__(256) k_adjoint_finalize(DevScene S, Wave w) {
    const WaveDesc d = cur_wave(w);
    for (uint32_t slot = blockIdx.x * blockDim.x + threadIdx.x; slot < d.count; slot += gridDim.x * blockDim.x) {
        f3 Af = f3_of(w.st_L[slot]) + f3_of(w.st_Le[slot]) + f3_of(w.st_La[slot]);
        if (!isfinite(average(Af))) continue;
        uint32_t tag = w.st_pix[slot];
        const SceneRec& R = S.recs[tag >> kSlotBits];
        const float* dl = S.adj_image + R.fb_off + 3ull * (tag & kSlotMask);
        f3 delta = mk3(dl[0], dl[1], dl[2]) * R.adj_scale;
        int depth = __float_as_int(w.st_prev[slot].w);
        for (int k = 1; k < depth && k <= (int)w.ad_depth; ++k) {
            size_t rec = (size_t)(k - 1) * w.capacity + slot;
            int2 t = w.ad_tex[rec];
            if (t.x < 0 && t.y < 0) continue;
            f3 Sk = Af - f3_of(w.ad_A[rec]);
            float4 a = w.ad_dlog[rec * 3], b = w.ad_dlog[rec * 3 + 1], c = w.ad_dlog[rec * 3 + 2];
            float wx = delta.x * Sk.x, wy = delta.y * Sk.y, wz = delta.z * Sk.z;
            float g0 = wx * a.x + wy * b.x + wz * c.x;
            float g1 = wx * a.y + wy * b.y + wz * c.y;
            float g2 = wx * a.z + wy * b.z + wz * c.z;
            float g3 = wx * a.w + wy * b.w + wz * c.w;
            if (t.x >= 0) {
                atomicAdd(&S.grad[t.x], g0);
                atomicAdd(&S.grad[t.x + 1], g1);
                atomicAdd(&S.grad[t.x + 2], g2);
            }
            if (t.y >= 0) atomicAdd(&S.grad[t.y], g3);
        }
    }
}

// FrameBuffer running mean of one scene (SeeSharp/Images/FrameBuffer.cs:89-160):
// at iteration n the image is scaled by (n-1)/n and the finite estimate / n added.
// The wave holds whole iterations, slot = iteration * W*H + pixel.
// FrameBuffer running mean over the wave's iterations (FrameBuffer.cs:89-160),
// and of the albedo / normal layers when they are written (RgbLayer.Splat,
// Layer.OnStartIteration: the same normalisation, no NaN filter).
// OutlierReplayCache (SeeSharp/Integrators/Util/OutlierReplayCache.cs:9-22): per
// pixel a .NET PriorityQueue<_, float> of at most n <= 4 samples keyed by the
// sample's average estimate, fed in iteration order. The queue is a quaternary
// min-heap (System.Collections.Generic.PriorityQueue, .NET 6+): Enqueue places
// the new entry at the end and moves it above its parent while strictly
// smaller; EnqueueDequeue (when full) replaces the root only if the new key is
// strictly larger, then moves down into the first strictly smallest child
// while that child is strictly smaller. With n <= 4 every entry is the root or
// its child. An entry is (weight bits << 32) | (iteration + 1), 0 = empty; the
// array is kept in heap order, which decides ties on read-out.
struct OutlierHeap {
    float w[4];
    uint32_t it[4];  // iteration + 1, 0 = empty
    int size;

    __device__ __forceinline__ void load(const unsigned long long* p, int n) {
        size = 0;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            unsigned long long e = j < n ? p[j] : 0ull;
            w[j] = __uint_as_float((uint32_t)(e >> 32));
            it[j] = (uint32_t)e;
            size += it[j] != 0u;
        }
    }
    __device__ __forceinline__ void store(unsigned long long* p, int n) const {
#pragma unroll
        for (int j = 0; j < 4; ++j)
            if (j < n) p[j] = ((unsigned long long)__float_as_uint(w[j]) << 32) | it[j];
    }
    __device__ __forceinline__ void set(int j, float wv, uint32_t iv) {
#pragma unroll
        for (int k = 0; k < 4; ++k)
            if (k == j) {
                w[k] = wv;
                it[k] = iv;
            }
    }
    __device__ __forceinline__ void add(float weight, uint32_t iteration, int n) {
        const uint32_t tag = iteration + 1u;
        if (size < n) {  // Enqueue: move up (the parent of entries 1..3 is entry 0)
            int idx = size++;
            if (idx > 0 && weight < w[0]) {
                set(idx, w[0], it[0]);
                idx = 0;
            }
            set(idx, weight, tag);
        } else if (size > 0 && weight > w[0]) {  // EnqueueDequeue: move down from the root
            int mc = 1;
            float mw = size > 1 ? w[1] : 0.0f;
            if (size > 2 && w[2] < mw) {
                mc = 2;
                mw = w[2];
            }
            if (size > 3 && w[3] < mw) {
                mc = 3;
                mw = w[3];
            }
            if (size > 1 && weight > mw) {
                uint32_t mi = mc == 1 ? it[1] : mc == 2 ? it[2] : it[3];
                w[0] = mw;
                it[0] = mi;
                set(mc, weight, tag);
            } else {
                w[0] = weight;
                it[0] = tag;
            }
        }
    }
};

// FrameBuffer running-mean steps of one pixel (FrameBuffer.StartIteration's
// scale by (n-1)/n, Splat's L/n, SeeSharp/Images/FrameBuffer.cs:89-160,232-255)
// over `num` samples in iteration order: sample s is in wave slot slot0 + s * stride
// and belongs to iteration first + s.
__device__ __forceinline__ void mean_steps(const Wave& w, uint32_t pix, uint32_t slot0, uint32_t stride,
                                           uint32_t num, uint32_t first, float* image, float* albedo,
                                           float* normal, unsigned long long* outliers, int num_outliers,
                                           uint32_t* nans) {
    f3 img = mk3(image[3 * pix], image[3 * pix + 1], image[3 * pix + 2]);
    f3 alb = gray(0.0f), nrm = gray(0.0f);
    if (albedo) {
        alb = mk3(albedo[3 * pix], albedo[3 * pix + 1], albedo[3 * pix + 2]);
        nrm = mk3(normal[3 * pix], normal[3 * pix + 1], normal[3 * pix + 2]);
    }
    OutlierHeap heap;
    unsigned long long* ok = outliers ? outliers + (size_t)pix * num_outliers : nullptr;
    if (ok) heap.load(ok, num_outliers);
    for (uint32_t s = 0; s < num; ++s) {
        uint32_t slot = slot0 + s * stride;
        f3 L = f3_of(w.st_L[slot]) + f3_of(w.st_Le[slot]) + f3_of(w.st_La[slot]);
        float n = (float)(first + s + 1);
        if (first + s + 1 > 1) img = img * ((n - 1.0f) / n);
        if (isfinite(average(L))) {
            img = img + div_exact(L, n);
            if (ok) heap.add(average(L), first + s, num_outliers);
        } else {
            ++*nans;
        }
        if (albedo) {
            if (first + s + 1 > 1) {
                alb = alb * ((n - 1.0f) / n);
                nrm = nrm * ((n - 1.0f) / n);
            }
            alb = alb + div_exact(f3_of(w.st_alb[slot]), n);
            nrm = nrm + div_exact(f3_of(w.st_nrm[slot]), n);
        }
    }
    image[3 * pix] = img.x;
    image[3 * pix + 1] = img.y;
    image[3 * pix + 2] = img.z;
    if (ok) heap.store(ok, num_outliers);
    if (albedo) {
        albedo[3 * pix] = alb.x;
        albedo[3 * pix + 1] = alb.y;
        albedo[3 * pix + 2] = alb.z;
        normal[3 * pix] = nrm.x;
        normal[3 * pix + 1] = nrm.y;
        normal[3 * pix + 2] = nrm.z;
    }
}

__global__ void __launch_bounds__(256) k_accumulate_mean(Wave w, float* image, uint32_t num_pixels, float* albedo,
                                                          float* normal, unsigned long long* outliers,
                                                          int num_outliers) {
    const WaveDesc d = cur_wave(w);
    uint32_t nans = 0;
    if (d.iters == 0) {
        // a wave over part of one or two iterations: render paths start .. start + count - 1
        // with count <= num_pixels, so no pixel occurs twice and each slot takes one step
        for (uint32_t slot = blockIdx.x * blockDim.x + threadIdx.x; slot < d.count;
             slot += gridDim.x * blockDim.x) {
            const uint32_t pix = w.st_pix[slot] & kSlotMask;
            const uint32_t it = d.first_iter + (uint32_t)((d.start + slot) / num_pixels);
            mean_steps(w, pix, slot, 0, 1, it, image, albedo, normal, outliers, num_outliers, &nans);
        }
    } else {
        // k: a path's index within its iteration; slot k of the wave's first
        // iteration records which pixel that is, and every iteration of the wave
        // lays its pixels out in the same order
        for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < num_pixels; k += gridDim.x * blockDim.x)
            mean_steps(w, w.st_pix[k] & kSlotMask, k, num_pixels, d.iters, d.first_iter, image, albedo, normal,
                       outliers, num_outliers, &nans);
    }
    if (nans) atomicAdd(&w.stats[(blockIdx.x & (kStatStripes - 1)) * kStatStride + kStatNan], (unsigned long long)nans);
}

// Descriptors of n running-mean waves of `per_wave` paths each (the last may be
// shorter) from render path first_path on, of a single-scene render of `total`
// paths starting at iteration first_iter: whole iterations when `whole` (the
// pixel count) divides per_wave, else (whole = 0) partial-iteration waves
// (iters = 0) of at most one iteration's paths.
__global__ void k_fill_waves(WaveDesc* descs, uint32_t n, uint64_t first_path, uint32_t per_wave, uint64_t total,
                             uint32_t first_iter, uint32_t whole) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const uint64_t start = first_path + (uint64_t)i * per_wave;
        WaveDesc d;
        d.start = start;
        d.count = (uint32_t)min((uint64_t)per_wave, total - start);
        d.first_iter = whole ? first_iter + (uint32_t)(start / whole) : first_iter;
        d.iters = whole ? d.count / whole : 0;
        d.pad = 0;
        descs[i] = d;
    }
}

// Batch accumulation: every finite path estimate is added to its scene's sum
// image (divided by the scene's sample count after the last wave).
__global__ void __launch_bounds__(256) k_accumulate_sum(DevScene S, Wave w, float* pool) {
    const WaveDesc d = cur_wave(w);
    for (uint32_t slot = blockIdx.x * blockDim.x + threadIdx.x; slot < d.count; slot += gridDim.x * blockDim.x) {
        f3 L = f3_of(w.st_L[slot]) + f3_of(w.st_Le[slot]) + f3_of(w.st_La[slot]);
        uint32_t tag = w.st_pix[slot];
        const SceneRec& R = S.recs[tag >> kSlotBits];
        if (S.outliers) {
            // A scene's paths are laid out iteration after iteration, so the
            // samples of one pixel in this wave sit num_pixels slots apart. The
            // slot holding the pixel's first of them feeds the pixel's queue
            // with all of them in iteration order (as the reference's
            // iterations are) and no other thread touches that queue.
            const uint64_t k = d.start + slot - R.path_off;  // the path's index within its scene
            const uint32_t np = R.num_pixels;
            if (k < np || slot < np) {
                OutlierHeap heap;
                unsigned long long* q = S.outliers + (R.fb_off / 3 + (tag & kSlotMask)) * (uint64_t)S.num_outliers;
                heap.load(q, S.num_outliers);
                const uint64_t scene_end = S.path_offsets[(tag >> kSlotBits) + 1] - d.start;  // in wave slots
                const uint32_t it0 = R.first_iter + (uint32_t)(k / np);
                uint32_t j = 0;
                for (uint64_t t = slot; t < d.count && t < scene_end; t += np, ++j) {
                    f3 Lt = f3_of(w.st_L[t]) + f3_of(w.st_Le[t]) + f3_of(w.st_La[t]);
                    if (isfinite(average(Lt))) heap.add(average(Lt), it0 + j, S.num_outliers);
                }
                heap.store(q, S.num_outliers);
            }
        }
        if (R.enable_denoiser) {
            float* a = S.layers + R.fb_off + 3ull * (tag & kSlotMask);
            float* nm = a + S.layer_stride;
            float4 av = w.st_alb[slot], nv = w.st_nrm[slot];
            atomicAdd(a, av.x);
            atomicAdd(a + 1, av.y);
            atomicAdd(a + 2, av.z);
            atomicAdd(nm, nv.x);
            atomicAdd(nm + 1, nv.y);
            atomicAdd(nm + 2, nv.z);
        }
        if (!isfinite(average(L))) {
            atomicAdd(&w.stats[stat_stripe() * kStatStride + kStatNan], 1ull);
            continue;
        }
        float* px = pool + R.fb_off + 3ull * (tag & kSlotMask);
        atomicAdd(px, L.x);
        atomicAdd(px + 1, L.y);
        atomicAdd(px + 2, L.z);
    }
}

__global__ void k_scale(float* data, uint64_t n, float factor) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        data[i] *= factor;
}

__global__ void __launch_bounds__(kTraceBlock) k_trace_closest(DevScene S, int scene, const float* rays, uint64_t n,
                                                               float4* out) {
    __shared__ int stack_node[kStackNodeLds];
    __shared__ float stack_t[kStackLds * kTraceBlock];
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float* r = rays + i * 8;
    const SceneRec& R = S.recs[scene];
    HitOut h = traverse<false>(S, R.root, R.abs_max, mk3(r[0], r[1], r[2]), mk3(r[4], r[5], r[6]), r[3], r[7],
                               stack_node + threadIdx.x, stack_t + threadIdx.x);
    out[i] = h.li >= 0 ? make_float4(h.t, h.u, h.v, __int_as_float(h.gid))
                       : make_float4(INFINITY, 0.0f, 0.0f, __int_as_float(-1));
}

__global__ void __launch_bounds__(kTraceBlock) k_trace_occluded(DevScene S, int scene, const float* rays,
                                                                uint64_t n, uint8_t* out) {
    __shared__ int stack_node[kStackNodeLds];
    __shared__ float stack_t[kStackLds * kTraceBlock];
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float* r = rays + i * 8;
    const SceneRec& R = S.recs[scene];
    HitOut h = traverse<true>(S, R.root, R.abs_max, mk3(r[0], r[1], r[2]), mk3(r[4], r[5], r[6]), r[3], r[7],
                              stack_node + threadIdx.x, stack_t + threadIdx.x);
    out[i] = h.li >= 0 ? 1 : 0;
}

__global__ void __launch_bounds__(256) k_primary_rays(DevScene S, int scene, uint32_t num_pixels, uint32_t base_seed,
                                                      uint32_t iteration, float* rays) {
    uint32_t pix = blockIdx.x * blockDim.x + threadIdx.x;
    if (pix >= num_pixels) return;
    const DevCamera& c = S.recs[scene].cam;
    uint32_t row = pix / (uint32_t)c.width, col = pix - row * (uint32_t)c.width;
    Rng rng{hash_seed(base_seed, pix, iteration)};
    f2 off = rng.next_float2();
    f3 o, d;
    camera_ray(c, (float)col + off.x, (float)row + off.y, &o, &d);
    float* r = rays + (uint64_t)pix * 8;
    r[0] = o.x, r[1] = o.y, r[2] = o.z, r[3] = 0.0f;
    r[4] = d.x, r[5] = d.y, r[6] = d.z, r[7] = INFINITY;
}

// ------------------------------------------------------------------ launch --
static inline uint32_t blocks(uint64_t n, uint32_t b) { return (uint32_t)((n + b - 1) / b); }

void launch_wave_begin(hipStream_t s, const Wave& w) { hipLaunchKernelGGL(k_wave_begin, dim3(1), dim3(1), 0, s, w); }
void launch_wave_end(hipStream_t s, const Wave& w) { hipLaunchKernelGGL(k_wave_end, dim3(1), dim3(1), 0, s, w); }
void launch_bounce_reset(hipStream_t s, const Wave& w, int buf) {
    hipLaunchKernelGGL(k_bounce_reset, dim3(1), dim3(1), 0, s, w, buf ^ 1);
}
static uint32_t grid_of(uint32_t configured, uint64_t max_items, uint32_t block) {
    if (configured) return configured;
    return (uint32_t)std::max<uint64_t>(1, (max_items + block - 1) / block);
}

// The prebuilt block sizes (the autotuner's block-shape axis): traversal kernels
// of 64, 128 (default) or 256 threads, forward shade blocks of 128, 256 (default)
// or 512 threads. A block's lanes share nothing but LDS stack space and queue
// appends, so every shape renders the same images.
template <class F>
static void with_trace_block(uint32_t tb, F&& f) {
    if (tb == 64) f(std::integral_constant<int, 64>{});
    else if (tb == 256) f(std::integral_constant<int, 256>{});
    else f(std::integral_constant<int, kTraceBlock>{});
}
template <class F>
static void with_shade_block(uint32_t sb, F&& f) {
    if (sb == 128) f(std::integral_constant<int, 128>{});
    else if (sb == 512) f(std::integral_constant<int, 512>{});
    else f(std::integral_constant<int, kShadeBlock>{});
}
bool valid_block_shape(uint32_t trace_block, uint32_t shade_block) {
    return (trace_block == 0 || trace_block == 64 || trace_block == 128 || trace_block == 256) &&
           (shade_block == 0 || shade_block == 128 || shade_block == 256 || shade_block == 512);
}

void launch_path(hipStream_t s, const DevScene& sc, const Wave& w, const LaunchConfig& lc) {
    hipLaunchKernelGGL(k_path, dim3(grid_of(lc.trace_blocks, lc.max_paths, kTraceBlock)), dim3(kTraceBlock), 0, s, sc,
                       w);
}

void launch_generate(hipStream_t s, const DevScene& sc, const Wave& w, const LaunchConfig& lc) {
    hipLaunchKernelGGL(k_generate, dim3(grid_of(lc.shade_blocks, lc.max_paths, 256)), dim3(256), 0, s, sc, w);
}
void launch_extend_primary(hipStream_t s, const DevScene& sc, const Wave& w, const LaunchConfig& lc) {
    with_trace_block(lc.trace_block, [&](auto tb) {
        constexpr int TB = decltype(tb)::value;
        hipLaunchKernelGGL(k_extend_primary<TB>, dim3(grid_of(lc.trace_blocks, lc.max_paths, TB)), dim3(TB), 0, s, sc,
                           w);
    });
}
void launch_extend(hipStream_t s, const DevScene& sc, const Wave& w, const LaunchConfig& lc, int buf) {
    with_trace_block(lc.trace_block, [&](auto tb) {
        constexpr int TB = decltype(tb)::value;
        hipLaunchKernelGGL(k_extend<TB>, dim3(grid_of(lc.trace_blocks, lc.max_paths, TB)), dim3(TB), 0, s, sc, w, buf);
    });
}
void launch_shade(hipStream_t s, const DevScene& sc, const Wave& w, const LaunchConfig& lc, int buf) {
    with_shade_block(lc.shade_block, [&](auto sbk) {
        constexpr int SB = decltype(sbk)::value;
        if (sc.num_scenes == 1 || sc.recs_alike)
            hipLaunchKernelGGL((k_shade<false, SB, true>), dim3(grid_of(lc.shade_blocks, lc.max_paths, SB)), dim3(SB),
                               0, s, sc, w, buf, lc.max_nsr);
        else
            hipLaunchKernelGGL((k_shade<false, SB, false>), dim3(grid_of(lc.shade_blocks, lc.max_paths, SB)), dim3(SB),
                               0, s, sc, w, buf, lc.max_nsr);
    });
}
void launch_shadow(hipStream_t s, const DevScene& sc, const Wave& w, const LaunchConfig& lc, int slot) {
    with_trace_block(lc.trace_block, [&](auto tb) {
        constexpr int TB = decltype(tb)::value;
        hipLaunchKernelGGL((k_shadow<false, TB>), dim3(grid_of(lc.trace_blocks, lc.max_shadow, TB)), dim3(TB), 0, s, sc,
                           w, slot);
    });
}
void launch_shade_adjoint(hipStream_t s, const DevScene& sc, const Wave& w, const LaunchConfig& lc, int buf) {
    // the adjoint shade keeps its one block size (its register budget is tuned for it)
    if (sc.num_scenes == 1 || sc.recs_alike)
        hipLaunchKernelGGL((k_shade<true, kShadeBlock, true>), dim3(grid_of(lc.shade_blocks, lc.max_paths, kShadeBlock)),
                           dim3(kShadeBlock), 0, s, sc, w, buf, lc.max_nsr);
    else
        hipLaunchKernelGGL((k_shade<true, kShadeBlock, false>), dim3(grid_of(lc.shade_blocks, lc.max_paths, kShadeBlock)),
                           dim3(kShadeBlock), 0, s, sc, w, buf, lc.max_nsr);
}
void launch_shade_grad(hipStream_t s, const DevScene& sc, const Wave& w, const LaunchConfig& lc, int buf) {
    hipLaunchKernelGGL(k_shade_grad, dim3(grid_of(lc.shade_blocks, (uint64_t)lc.max_paths + lc.max_shadow, 256)),
                       dim3(256), 0, s, sc, w, buf);
}
void launch_trace(hipStream_t s, const DevScene& sc, const Wave& w, const LaunchConfig& lc, int buf, bool adjoint) {
    with_trace_block(lc.trace_block, [&](auto tb) {
        constexpr int TB = decltype(tb)::value;
        const uint32_t ne = grid_of(lc.trace_blocks, lc.max_paths, TB);
        const uint32_t ns = grid_of(lc.trace_blocks, lc.max_shadow, TB);
        if (adjoint)
            hipLaunchKernelGGL((k_trace<true, TB>), dim3(ne + ns), dim3(TB), 0, s, sc, w, buf, ne);
        else
            hipLaunchKernelGGL((k_trace<false, TB>), dim3(ne + ns), dim3(TB), 0, s, sc, w, buf, ne);
    });
}
void launch_shadow_adjoint(hipStream_t s, const DevScene& sc, const Wave& w, const LaunchConfig& lc, int slot) {
    with_trace_block(lc.trace_block, [&](auto tb) {
        constexpr int TB = decltype(tb)::value;
        hipLaunchKernelGGL((k_shadow<true, TB>), dim3(grid_of(lc.trace_blocks, lc.max_shadow, TB)), dim3(TB), 0, s, sc,
                           w, slot);
    });
}
void launch_adjoint_finalize(hipStream_t s, const DevScene& sc, const Wave& w, const LaunchConfig& lc) {
    hipLaunchKernelGGL(k_adjoint_finalize, dim3(grid_of(lc.shade_blocks, lc.max_paths, 256)), dim3(256), 0, s, sc, w);
}
void launch_accumulate_mean(hipStream_t s, const Wave& w, float* image, uint32_t num_pixels, float* albedo,
                            float* normal, unsigned long long* outliers, int num_outliers) {
    hipLaunchKernelGGL(k_accumulate_mean, dim3(std::min(blocks(num_pixels, 256), 8192u)), dim3(256), 0, s, w, image,
                       num_pixels, albedo, normal, outliers, num_outliers);
}
void launch_fill_waves(hipStream_t s, WaveDesc* descs, uint32_t n, uint64_t first_path, uint32_t per_wave,
                       uint64_t total, uint32_t first_iter, uint32_t whole) {
    hipLaunchKernelGGL(k_fill_waves, dim3(std::max(1u, std::min(blocks(n, 256), 64u))), dim3(256), 0, s, descs, n,
                       first_path, per_wave, total, first_iter, whole);
}

void launch_accumulate_sum(hipStream_t s, const DevScene& sc, const Wave& w, const LaunchConfig& lc, float* pool) {
    hipLaunchKernelGGL(k_accumulate_sum, dim3(grid_of(lc.shade_blocks, lc.max_paths, 256)), dim3(256), 0, s, sc, w,
                       pool);
}
void resident_blocks_per_cu(uint32_t trace_block, uint32_t shade_block, int* trace, int* shade) {
    int e = 0, sh = 0, sd = 0;
    with_trace_block(trace_block, [&](auto tb) {
        constexpr int TB = decltype(tb)::value;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&e, k_extend<TB>, TB, 0) != hipSuccess) e = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&sh, k_shadow<false, TB>, TB, 0) != hipSuccess) sh = 0;
    });
    with_shade_block(shade_block, [&](auto sbk) {
        constexpr int SB = decltype(sbk)::value;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&sd, k_shade<false, SB, false>, SB, 0) != hipSuccess) sd = 0;
    });
    *trace = std::max(e, sh);
    *shade = sd;
}

void launch_scale(hipStream_t s, float* data, uint64_t n, float factor) {
    if (!n) return;
    hipLaunchKernelGGL(k_scale, dim3((uint32_t)std::min<uint64_t>(blocks(n, 256), 4096)), dim3(256), 0, s, data, n,
                       factor);
}
void launch_trace_closest(hipStream_t s, const DevScene& sc, int scene, const float* rays, uint64_t n, float4* out) {
    if (!n) return;
    hipLaunchKernelGGL(k_trace_closest, dim3(blocks(n, kTraceBlock)), dim3(kTraceBlock), 0, s, sc, scene, rays, n,
                       out);
}
void launch_trace_occluded(hipStream_t s, const DevScene& sc, int scene, const float* rays, uint64_t n,
                           uint8_t* out) {
    if (!n) return;
    hipLaunchKernelGGL(k_trace_occluded, dim3(blocks(n, kTraceBlock)), dim3(kTraceBlock), 0, s, sc, scene, rays, n,
                       out);
}
void launch_primary_rays(hipStream_t s, const DevScene& sc, int scene, uint32_t num_pixels, uint32_t base_seed,
                         uint32_t iteration, float* rays) {
    hipLaunchKernelGGL(k_primary_rays, dim3(blocks(num_pixels, 256)), dim3(256), 0, s, sc, scene, num_pixels,
                       base_seed, iteration, rays);
}

void launch_light_paths(hipStream_t s, const DevScene& sc, int scene, const LightParams& p, float* image,
                        unsigned long long* stats) {
    hipLaunchKernelGGL(k_light_paths, dim3(blocks(p.num_paths, kTraceBlock)), dim3(kTraceBlock), 0, s, sc, scene, p,
                       image, stats);
}

void launch_bidir_light(hipStream_t s, const DevScene& sc, int scene, const BidirParams& p, float* image,
                        unsigned long long* stats) {
    hipLaunchKernelGGL(p.env ? k_bidir_light<true> : k_bidir_light<false>, dim3(blocks(p.num_paths, kTraceBlock)),
                       dim3(kTraceBlock), 0, s, sc, scene, p, image, stats);
}

void launch_bidir_camera_wave(hipStream_t s, const DevScene& sc, int scene, const BidirParams& p, float* image,
                              unsigned long long* stats) {
    const uint32_t n = p.num_pixels;
    const uint32_t grid = blocks(n, 256);
    hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(1), 0, s, p.bc_count, (uint64_t)1, n);
    hipLaunchKernelGGL(p.env ? k_bc_init<true> : k_bc_init<false>, dim3(grid), dim3(256), 0, s, sc, scene, p);
    int in = 0;
    for (int depth = 1; depth < p.max_depth + 1; ++depth, in = 1 - in) {
        // the next bounce's list and the record list start empty
        hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(1), 0, s, p.bc_count + (1 - in), (uint64_t)1, 0u);
        hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(1), 0, s, p.bc_count + 2, (uint64_t)1, 0u);
        hipLaunchKernelGGL(k_bc_trace, dim3(blocks(n, kTraceBlock)), dim3(kTraceBlock), 0, s, sc, scene, p, in, stats);
        hipLaunchKernelGGL(p.env ? k_bc_shade<true> : k_bc_shade<false>, dim3(blocks(n, kTraceBlock)),
                           dim3(kTraceBlock), 0, s, sc, scene, p, in, stats);
        hipLaunchKernelGGL(k_bc_shadow, dim3(blocks((uint64_t)n * std::max(p.bc_rec, 1), kTraceBlock)),
                           dim3(kTraceBlock), 0, s, sc, scene, p, stats);
        hipLaunchKernelGGL(k_bc_gather, dim3(grid), dim3(256), 0, s, p, in);
    }
    hipLaunchKernelGGL(k_bc_finish, dim3(grid), dim3(256), 0, s, p, image, stats);
}

void launch_vcm_prepare(hipStream_t s, const DevScene& sc, int scene, const BidirParams& p, int* scan_tmp) {
    (void)scene;
    auto scan = [&](const int* in, int* out, uint32_t n) {
        if (n == 0) return;
        const uint32_t nb = (n + kScanBlock - 1) / kScanBlock;
        hipLaunchKernelGGL(k_vscan_local, dim3(nb), dim3(kScanBlock), 0, s, in, out, n, scan_tmp);
        hipLaunchKernelGGL(k_vscan_carry, dim3(1), dim3(kScanBlock), 0, s, scan_tmp, nb);
        hipLaunchKernelGGL(k_vscan_add, dim3(nb), dim3(kScanBlock), 0, s, out, n, scan_tmp);
    };
    scan(p.lv_len, p.cum, p.num_paths);  // PathCache.Prepare (PathCache.cs:87-94)
    if (p.mode != kModeVcm || !p.merging || p.max_depth < 2) return;
    const uint32_t table = p.table_mask + 1;
    const uint64_t photons = (uint64_t)(p.max_depth - 1) * p.num_paths;
    if (photons == 0) return;
    const uint32_t grid = (uint32_t)std::min<uint64_t>(blocks(photons, 256), 65536);
    hipLaunchKernelGGL(k_fill_u32, dim3(blocks(table, 256)), dim3(256), 0, s, p.bucket_fill, (uint64_t)table, 0u);
    hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(1), 0, s, p.bucket_start, (uint64_t)1, 0u);
    hipLaunchKernelGGL(k_photon_grid<false>, dim3(grid), dim3(256), 0, s, sc, p);
    scan(reinterpret_cast<const int*>(p.bucket_fill), reinterpret_cast<int*>(p.bucket_start + 1), table);
    hipLaunchKernelGGL(k_fill_u32, dim3(blocks(table, 256)), dim3(256), 0, s, p.bucket_fill, (uint64_t)table, 0u);
    hipLaunchKernelGGL(k_photon_grid<true>, dim3(grid), dim3(256), 0, s, sc, p);
}

void launch_vcm_camera(hipStream_t s, const DevScene& sc, int scene, const BidirParams& p, float* image,
                       unsigned long long* stats) {
    if (p.light_tracer && p.num_paths > 0)
        hipLaunchKernelGGL(p.env ? k_vcm_splat<true> : k_vcm_splat<false>, dim3(blocks(p.num_paths, kTraceBlock)),
                           dim3(kTraceBlock), 0, s, sc, scene, p, image, stats);
    hipLaunchKernelGGL(p.env ? k_vcm_camera<true> : k_vcm_camera<false>, dim3(blocks(p.num_pixels, kTraceBlock)),
                       dim3(kTraceBlock), 0, s, sc, scene, p, image, stats);
}

void launch_debug_visualize(hipStream_t s, const DevScene& sc, int scene, uint32_t num_pixels, uint32_t base_seed,
                            uint32_t first, uint32_t num, float* image) {
    hipLaunchKernelGGL(k_debug_visualize, dim3(blocks(num_pixels, kTraceBlock)), dim3(kTraceBlock), 0, s, sc, scene,
                       base_seed, first, num, image);
}

void launch_replay(hipStream_t s, const DevScene& sc, int scene, uint32_t pix, uint32_t iteration, PathNode* nodes,
                   uint32_t cap, uint32_t* count, float* estimate) {
    hipLaunchKernelGGL(k_replay, dim3(1), dim3(kTraceBlock), 0, s, sc, scene, pix, iteration, nodes, cap, count,
                       estimate);
}

}  // namespace sspt
