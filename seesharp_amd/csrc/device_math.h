// device_math.h -- fp32 vector math, SeeSharp's RNG and sampling warps as
// device functions for gfx950.
//
// Arithmetic rules shared with the host side of the library (the whole
// library is compiled with -ffp-contract=off so that camera rays, the RNG
// jitter and the watertight triangle test round exactly like the CPU):
//   * vector sums are evaluated left to right, x*x + y*y + z*z;
//   * normalize_exact(v) = v / sqrt(dot(v, v)) per component (correctly
//     rounded division and square root), for camera rays; shading uses the
//     hardware forms (see ssqrt);
//   * (int)float follows C#'s x86 conversion: truncation, out-of-range and
//     NaN give INT_MIN.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define SSPT_HD __host__ __device__ __forceinline__

namespace sspt {

constexpr float kPi = 3.14159265358979323846f;
constexpr float kInvPi = 0.318309886183790671538f;
constexpr float kInv2Pi = 0.159154943091895335769f;

struct f3 {
    float x, y, z;
};
struct f2 {
    float x, y;
};

SSPT_HD f3 mk3(float x, float y, float z) { return f3{x, y, z}; }
SSPT_HD f3 operator+(f3 a, f3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
SSPT_HD f3 operator-(f3 a, f3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
SSPT_HD f3 operator-(f3 a) { return {-a.x, -a.y, -a.z}; }
SSPT_HD f3 operator*(f3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
SSPT_HD f3 operator*(float s, f3 a) { return {s * a.x, s * a.y, s * a.z}; }
SSPT_HD f3 operator*(f3 a, f3 b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
// Shading arithmetic on the device uses the hardware square root and
// reciprocal (within about 1 ulp) in place of the correctly rounded sequences
// of a dozen instructions each: ssqrt, srcp, length, normalize and f3 / s.
// Host code keeps the exact forms. Camera rays use normalize_exact and the
// triangle test exact scalar division, so both still round like the CPU.
SSPT_HD float ssqrt(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_sqrtf(x);
#else
    return sqrtf(x);
#endif
}
SSPT_HD float srcp(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_rcpf(x);
#else
    return 1.0f / x;
#endif
}
SSPT_HD float sdiv(float a, float b) {
#if defined(__HIP_DEVICE_COMPILE__)
    return a * srcp(b);
#else
    return a / b;
#endif
}
// sine and cosine of an angle in [0, 2 pi] (hardware forms on the device,
// absolute error about 1e-6)
SSPT_HD float ssin(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __sinf(x);
#else
    return sinf(x);
#endif
}
SSPT_HD void ssincos(float x, float* s, float* c) {
#if defined(__HIP_DEVICE_COMPILE__)
    *s = __sinf(x);
    *c = __cosf(x);
#else
    *s = sinf(x);
    *c = cosf(x);
#endif
}
SSPT_HD f3 operator/(f3 a, float s) {
#if defined(__HIP_DEVICE_COMPILE__)
    const float r = srcp(s);
    return {a.x * r, a.y * r, a.z * r};
#else
    return {a.x / s, a.y / s, a.z / s};
#endif
}
SSPT_HD bool is_zero(f3 a) { return a.x == 0.0f && a.y == 0.0f && a.z == 0.0f; }
SSPT_HD float dot(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
SSPT_HD f3 cross(f3 a, f3 b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
SSPT_HD float length(f3 a) { return ssqrt(dot(a, a)); }
SSPT_HD f3 normalize(f3 a) { return a / length(a); }
// correctly rounded per-component division (FrameBuffer arithmetic)
SSPT_HD f3 div_exact(f3 a, float s) { return {a.x / s, a.y / s, a.z / s}; }
// correctly rounded: v / sqrt(dot(v, v)) per component
SSPT_HD f3 normalize_exact(f3 a) {
    const float l = sqrtf(dot(a, a));
    return {a.x / l, a.y / l, a.z / l};
}
SSPT_HD f3 vabs(f3 a) { return {fabsf(a.x), fabsf(a.y), fabsf(a.z)}; }
SSPT_HD float comp(f3 a, int k) { return k == 0 ? a.x : (k == 1 ? a.y : a.z); }
SSPT_HD f3 f3_of(float4 v) { return {v.x, v.y, v.z}; }
SSPT_HD float4 f4_of(f3 v, float w) { return make_float4(v.x, v.y, v.z, w); }

// RgbColor (SimpleImageIO)
SSPT_HD f3 gray(float v) { return {v, v, v}; }
SSPT_HD float average(f3 c) { return (c.x + c.y + c.z) / 3.0f; }
SSPT_HD float luminance(f3 c) { return 0.2126f * c.x + 0.7152f * c.y + 0.0722f * c.z; }
SSPT_HD f3 lerp(float w, f3 a, f3 b) { return (1.0f - w) * a + w * b; }

SSPT_HD int to_int(float f) {
    if (!(f > -2147483904.0f && f < 2147483648.0f)) return (int)0x80000000;
    return (int)f;
}
SSPT_HD float clampf(float v, float lo, float hi) { return fminf(fmaxf(v, lo), hi); }

// ------------------------------------------------------------------ RNG ---
// SeeSharp/Sampling/RNG.cs:90-118 (PCG-RXS-M-XS state, FNV seed mixing).
SSPT_HD uint32_t pcg_hash(uint32_t input) {
    uint32_t state = input * 747796405u + 2891336453u;
    uint32_t word = ((state >> ((state >> 28) + 4)) ^ state) * 277803737u;
    return (word >> 22) ^ word;
}
SSPT_HD uint32_t fnv_hash(uint32_t hash, uint32_t data) {
    hash = (hash * 16777619u) ^ (data & 0xFF);
    hash = (hash * 16777619u) ^ ((data >> 8) & 0xFF);
    hash = (hash * 16777619u) ^ ((data >> 16) & 0xFF);
    hash = (hash * 16777619u) ^ ((data >> 24) & 0xFF);
    return hash;
}
SSPT_HD uint32_t hash_seed(uint32_t base, uint32_t chain, uint32_t sample) {
    uint32_t h = fnv_hash(2166136261u, pcg_hash(base));
    h = fnv_hash(h, pcg_hash(chain));
    h = fnv_hash(h, pcg_hash(sample));
    return pcg_hash(h);
}

struct Rng {
    uint32_t state;
    SSPT_HD uint32_t next() {
        uint32_t word = ((state >> ((state >> 28) + 4)) ^ state) * 277803737u;
        state = state * 747796405u + 2891336453u;
        return (word >> 22) ^ word;
    }
    // Next() / (float)uint.MaxValue: u32 -> f32 rounds to nearest; the
    // divisor is 2^32 in fp32, so the division is an exact scaling.
    SSPT_HD float next_float() { return (float)next() * 2.3283064365386963e-10f; }
    SSPT_HD f2 next_float2() {
        float a = next_float();
        float b = next_float();
        return {a, b};
    }
    // RNG.NextInt(uint) (RNG.cs:56-70): Lemire's multiply-shift with rejection.
    SSPT_HD uint32_t next_int(uint32_t max) {
        uint32_t x = next();
        uint64_t m = (uint64_t)x * max;
        uint32_t l = (uint32_t)m;
        if (l < max) {
            uint32_t t = (0u - max) % max;
            while (l < t) {
                x = next();
                m = (uint64_t)x * max;
                l = (uint32_t)m;
            }
        }
        return (uint32_t)(m >> 32);
    }
};

// ---------------------------------------------------------------- warps ---
// SampleWarp.ToUniformTriangle (SeeSharp/Sampling/SampleWarp.cs:9-14)
SSPT_HD f2 to_uniform_triangle(f2 p) {
    float s = ssqrt(p.x);
    return {1.0f - s, p.y * s};
}
// SampleWarp.ToCosHemisphere + SphericalToCartesian (SampleWarp.cs:89-96,130-137)
SSPT_HD f3 to_cos_hemisphere(f2 p) {
    float sin_t = ssqrt(1.0f - p.y), cos_t = ssqrt(p.y);
    float phi = 2.0f * kPi * p.x, sp, cp;
    ssincos(phi, &sp, &cp);
    return {sin_t * cp, sin_t * sp, cos_t};
}
// SampleWarp.SurfaceAreaToSolidAngle (SampleWarp.cs:290-295)
SSPT_HD float area_to_solid_angle(f3 from, f3 to, f3 to_normal) {
    f3 d = to - from;
    float d2 = dot(d, d);
    return sdiv(fabsf(dot(to_normal, -d)), d2 * ssqrt(d2));
}

// ---------------------------------------------------------- shading space -
// ShadingSpace.ComputeBasisVectors / WorldToShading / ShadingToWorld
// (SeeSharp/Shading/ShadingSpace.cs:14-71)
struct Frame {
    f3 n, t, b;
    SSPT_HD void build(f3 normal) {
        n = normal;
        if (fabsf(n.x) > fabsf(n.y)) {
            float denom = ssqrt(n.x * n.x + n.z * n.z);
            t = mk3(-n.z, 0.0f, n.x) / denom;
        } else {
            float denom = ssqrt(n.y * n.y + n.z * n.z);
            t = mk3(0.0f, n.z, -n.y) / denom;
        }
        b = cross(n, t);
    }
    SSPT_HD f3 to_local(f3 d) const { return {dot(t, d), dot(b, d), dot(n, d)}; }
    SSPT_HD f3 to_world(f3 d) const { return d.z * n + d.x * t + d.y * b; }
};

SSPT_HD float cos2_theta(f3 d) { return d.z * d.z; }
SSPT_HD float sin2_theta(f3 d) { return fmaxf(0.0f, 1.0f - cos2_theta(d)); }
SSPT_HD float sin_theta(f3 d) { return ssqrt(sin2_theta(d)); }
SSPT_HD float tan_theta(f3 d) { return sdiv(sin_theta(d), d.z); }
SSPT_HD float tan2_theta(f3 d) { return sdiv(sin2_theta(d), cos2_theta(d)); }
SSPT_HD float cos_phi(f3 d) {
    float s = sin_theta(d);
    return s == 0.0f ? 1.0f : clampf(sdiv(d.x, s), -1.0f, 1.0f);
}
SSPT_HD float sin_phi(f3 d) {
    float s = sin_theta(d);
    return s == 0.0f ? 0.0f : clampf(sdiv(d.y, s), -1.0f, 1.0f);
}
SSPT_HD f3 reflect(f3 o, f3 n) { return -o + 2.0f * dot(o, n) * n; }
SSPT_HD bool refract(f3 i, f3 n, float eta, f3* out) {
    float cos_i = dot(n, i);
    float sin2_i = fmaxf(0.0f, 1.0f - cos_i * cos_i);
    float sin2_t = eta * eta * sin2_i;
    if (sin2_t >= 1.0f) return false;
    float cos_t = ssqrt(1.0f - sin2_t);
    *out = eta * -i + (eta * cos_i - cos_t) * n;
    return true;
}
SSPT_HD bool same_hemisphere(f3 a, f3 b) { return a.z * b.z > 0.0f; }

// ---------------------------------------------------------- microfacet ----
// TrowbridgeReitzDistribution (SeeSharp/Shading/TrowbridgeReitzDistribution.cs)
struct Ggx {
    float ax, ay;
    SSPT_HD float ndf(f3 n) const {
        float tan2 = tan2_theta(n);
        if (isinf(tan2)) return 0.0f;
        float cos4 = cos2_theta(n) * cos2_theta(n);
        float cp = cos_phi(n), sp = sin_phi(n);
        float e = tan2 * (sdiv(cp * cp, ax * ax) + sdiv(sp * sp, ay * ay));
        return srcp(kPi * ax * ay * cos4 * (1.0f + e) * (1.0f + e));
    }
    SSPT_HD float masking_ratio(f3 o) const {
        float abs_tan = fabsf(tan_theta(o));
        if (isinf(abs_tan)) return 0.0f;
        float cp = cos_phi(o), sp = sin_phi(o);
        float alpha = ssqrt(cp * cp * ax * ax + sp * sp * ay * ay);
        float a2t2 = alpha * abs_tan * (alpha * abs_tan);
        return (-1.0f + ssqrt(1.0f + a2t2)) * 0.5f;
    }
    SSPT_HD float g1(f3 o) const { return srcp(1.0f + masking_ratio(o)); }
    SSPT_HD float g(f3 o, f3 i) const { return srcp(1.0f + masking_ratio(o) + masking_ratio(i)); }
    SSPT_HD float pdf(f3 o, f3 n) const {
        if (!same_hemisphere(o, n)) return 0.0f;
        return sdiv(g1(o) * fmaxf(0.0f, dot(o, n)) * ndf(n), fabsf(o.z));
    }
    // Heitz 2018 visible-normal sampling (TrowbridgeReitzDistribution.cs:70-100);
    // the reciprocal square-root estimate is replaced by 1 / sqrt.
    SSPT_HD f3 sample(f3 o, f2 p) const {
        bool flip = false;
        if (o.z < 0.0f) {
            flip = true;
            o = -o;
        }
        f3 vh = normalize(mk3(ax * o.x, ay * o.y, o.z));
        float lensq = vh.x * vh.x + vh.y * vh.y;
        f3 t1 = lensq > 0.0f ? mk3(-vh.y, vh.x, 0.0f) * srcp(ssqrt(lensq)) : mk3(1, 0, 0);
        f3 t2 = cross(vh, t1);
        float r = ssqrt(p.x);
        float phi = 2.0f * kPi * p.y, sphi, cphi;
        ssincos(phi, &sphi, &cphi);
        float a1 = r * cphi;
        float a2 = r * sphi;
        float s = 0.5f * (1.0f + vh.z);
        a2 = (1.0f - s) * ssqrt(1.0f - a1 * a1) + s * a2;
        f3 nh = a1 * t1 + a2 * t2 + ssqrt(fmaxf(0.0f, 1.0f - a1 * a1 - a2 * a2)) * vh;
        f3 ne = normalize(mk3(ax * nh.x, ay * nh.y, fmaxf(0.0f, nh.z)));
        return flip ? -ne : ne;
    }
};

// Fresnel (SeeSharp/Shading/Fresnel.cs)
SSPT_HD float fresnel_dielectric(float cos_i, float eta_i, float eta_t) {
    cos_i = clampf(cos_i, -1.0f, 1.0f);
    if (!(cos_i > 0.0f)) {
        float tmp = eta_i;
        eta_i = eta_t;
        eta_t = tmp;
        cos_i = fabsf(cos_i);
    }
    float sin_i = ssqrt(fmaxf(0.0f, 1.0f - cos_i * cos_i));
    float sin_t = sdiv(eta_i, eta_t) * sin_i;
    if (sin_t >= 1.0f) return 1.0f;
    float cos_t = ssqrt(fmaxf(0.0f, 1.0f - sin_t * sin_t));
    float rparl = sdiv(eta_t * cos_i - eta_i * cos_t, eta_t * cos_i + eta_i * cos_t);
    float rperp = sdiv(eta_i * cos_i - eta_t * cos_t, eta_i * cos_i + eta_t * cos_t);
    return (rparl * rparl + rperp * rperp) * 0.5f;
}
SSPT_HD float schlick_weight(float c) {
    float m = clampf(1.0f - c, 0.0f, 1.0f);
    return (m * m) * (m * m) * m;
}
SSPT_HD f3 schlick_fresnel(f3 r0, float c) { return lerp(schlick_weight(c), r0, gray(1.0f)); }
SSPT_HD float schlick_r0(float eta) {
    float r = sdiv(eta - 1.0f, eta + 1.0f);
    return r * r;
}

}  // namespace sspt
