// adjoint.h -- derivatives of the material value with respect to the texture
// parameters of the shaded vertex, for the path-replay adjoint.
//
// Forward-mode dual numbers with four tangents: d/d(base color r, g, b) and
// d/d(roughness). Only the BSDF value is differentiated; sampling pdfs, MIS
// weights and the Russian-roulette probability are treated as constants
// ("detached" sampling), which keeps the gradient estimator unbiased.
#pragma once

#include "shading.h"

namespace sspt {

struct D4 {
    float v;
    float d[4];
};

__device__ __forceinline__ D4 dconst(float v) { return D4{v, {0.0f, 0.0f, 0.0f, 0.0f}}; }
__device__ __forceinline__ D4 dvar(float v, int k) {
    D4 r = dconst(v);
    r.d[k] = 1.0f;
    return r;
}
__device__ __forceinline__ D4 operator+(D4 a, D4 b) {
    D4 r;
    r.v = a.v + b.v;
#pragma unroll
    for (int k = 0; k < 4; ++k) r.d[k] = a.d[k] + b.d[k];
    return r;
}
__device__ __forceinline__ D4 operator-(D4 a, D4 b) {
    D4 r;
    r.v = a.v - b.v;
#pragma unroll
    for (int k = 0; k < 4; ++k) r.d[k] = a.d[k] - b.d[k];
    return r;
}
__device__ __forceinline__ D4 operator*(D4 a, D4 b) {
    D4 r;
    r.v = a.v * b.v;
#pragma unroll
    for (int k = 0; k < 4; ++k) r.d[k] = a.d[k] * b.v + a.v * b.d[k];
    return r;
}
__device__ __forceinline__ D4 operator/(D4 a, D4 b) {
    D4 r;
    const float ib = srcp(b.v);
    r.v = a.v * ib;
    const float inv2 = ib * ib;
#pragma unroll
    for (int k = 0; k < 4; ++k) r.d[k] = (a.d[k] * b.v - a.v * b.d[k]) * inv2;
    return r;
}
__device__ __forceinline__ D4 operator+(D4 a, float b) {
    a.v += b;
    return a;
}
__device__ __forceinline__ D4 operator+(float b, D4 a) { return a + b; }
__device__ __forceinline__ D4 operator-(D4 a, float b) {
    a.v -= b;
    return a;
}
__device__ __forceinline__ D4 operator-(float b, D4 a) {
    D4 r;
    r.v = b - a.v;
#pragma unroll
    for (int k = 0; k < 4; ++k) r.d[k] = -a.d[k];
    return r;
}
__device__ __forceinline__ D4 operator*(D4 a, float b) {
    a.v *= b;
#pragma unroll
    for (int k = 0; k < 4; ++k) a.d[k] *= b;
    return a;
}
__device__ __forceinline__ D4 operator*(float b, D4 a) { return a * b; }
__device__ __forceinline__ D4 operator/(D4 a, float b) {
    const float ib = srcp(b);
    a.v *= ib;
#pragma unroll
    for (int k = 0; k < 4; ++k) a.d[k] *= ib;
    return a;
}
__device__ __forceinline__ D4 operator/(float a, D4 b) { return dconst(a) / b; }
__device__ __forceinline__ D4 dsqrt(D4 a) {
    D4 r;
    r.v = ssqrt(a.v);
    float g = r.v > 0.0f ? 0.5f * srcp(r.v) : 0.0f;
#pragma unroll
    for (int k = 0; k < 4; ++k) r.d[k] = a.d[k] * g;
    return r;
}
// max(c, a): the derivative of the branch that is taken
__device__ __forceinline__ D4 dmax(float c, D4 a) { return a.v >= c ? a : dconst(c); }

struct D3 {
    D4 x, y, z;
};
__device__ __forceinline__ D3 d3const(f3 c) { return D3{dconst(c.x), dconst(c.y), dconst(c.z)}; }
__device__ __forceinline__ D3 operator+(D3 a, D3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ D3 operator*(D3 a, D4 s) { return {a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ D3 operator*(D3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ D3 operator*(D3 a, D3 b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
__device__ __forceinline__ D3 operator/(D3 a, D4 s) { return {a.x / s, a.y / s, a.z / s}; }
__device__ __forceinline__ D3 operator/(D3 a, float s) {
    const float is = srcp(s);
    return {a.x * is, a.y * is, a.z * is};
}
__device__ __forceinline__ D3 d3lerp(float w, D3 a, D3 b) { return a * (1.0f - w) + b * w; }
__device__ __forceinline__ D3 d3lerp(float w, f3 a, D3 b) { return d3const(a) * (1.0f - w) + b * w; }
__device__ __forceinline__ D3 d3lerp(float w, D3 a, f3 b) { return a * (1.0f - w) + d3const(b) * w; }

// Trowbridge-Reitz NDF and Smith masking (TrowbridgeReitzDistribution.cs:22-50)
// with dual alphas.
struct GgxD {
    D4 ax, ay;
    __device__ D4 ndf(f3 n) const {
        float tan2 = tan2_theta(n);
        if (isinf(tan2)) return dconst(0.0f);
        float cos4 = cos2_theta(n) * cos2_theta(n);
        float cp = cos_phi(n), sp = sin_phi(n);
        D4 e = tan2 * (cp * cp / (ax * ax) + sp * sp / (ay * ay));
        D4 ope = 1.0f + e;
        return 1.0f / (kPi * ax * ay * cos4 * ope * ope);
    }
    __device__ D4 masking_ratio(f3 o) const {
        float abs_tan = fabsf(tan_theta(o));
        if (isinf(abs_tan)) return dconst(0.0f);
        float cp = cos_phi(o), sp = sin_phi(o);
        D4 alpha = dsqrt(cp * cp * ax * ax + sp * sp * ay * ay);
        D4 at = alpha * abs_tan;
        D4 a2t2 = at * at;
        return (dsqrt(1.0f + a2t2) - 1.0f) * 0.5f;
    }
    __device__ D4 g(f3 o, f3 i) const { return 1.0f / (1.0f + masking_ratio(o) + masking_ratio(i)); }
};

// Value of the shaded vertex's material for the local direction `in`, with
// tangents. Mirrors generic_value_pdf / diffuse_value_pdf in shading.h.
// tex_base / tex_rough select which parameters carry tangents (textured ones).
__device__ D3 material_value_dual(const DevScene& S, const Shader& sh, f2 uv, f3 in, bool tex_base, bool tex_rough) {
    const DevMaterial& m = sh.m;
    f3 base_f = sh.base;
    D3 base{tex_base ? dvar(base_f.x, 0) : dconst(base_f.x), tex_base ? dvar(base_f.y, 1) : dconst(base_f.y),
            tex_base ? dvar(base_f.z, 2) : dconst(base_f.z)};
    f3 out = sh.out_local;
    bool same_geo = should_reflect(sh.ng, sh.out_world, sh.f.to_world(in));
    bool same_hemi = same_hemisphere(out, in);
    if (m.kind != 1) {
        // DiffuseMaterial.Evaluate (DiffuseMaterial.cs:127-139)
        bool ok = (m.transmitter && !same_geo) ? !same_hemi : (same_geo && same_hemi);
        return ok ? base / kPi : d3const(gray(0.0f));
    }
    // GenericMaterial.LocalParams (GenericMaterial.cs:354-381)
    D4 lum = base.x * 0.2126f + base.y * 0.7152f + base.z * 0.0722f;
    D3 color_tint = lum.v > 0.0f ? base / lum : d3const(gray(1.0f));
    D3 spec_tint = d3lerp(m.spec_tint, gray(1.0f), color_tint);
    D4 rough = tex_rough ? dvar(sh.roughness, 3) : dconst(sh.roughness);
    float aspect = ssqrt(1.0f - m.anisotropic * 0.9f);
    GgxD ggx{dmax(0.001f, rough * rough / aspect), dmax(0.001f, rough * rough * aspect)};
    D3 spec_r0 = d3lerp(m.metallic, spec_tint * schlick_r0(m.ior), base);
    float dw = (1.0f - m.metallic) * (1.0f - m.spec_trans);
    D3 diffuse_refl = base * dw;
    D3 spec_trans = base * m.spec_trans;

    // GenericMaterial.ComputeValueAndPdf value (GenericMaterial.cs:152-296)
    float cos_o = fabsf(out.z), cos_i = fabsf(in.z);
    D3 val = d3const(gray(0.0f));
    float f_out = schlick_weight(cos_o), f_in = schlick_weight(cos_i);
    if (same_hemi && same_geo) val = val + diffuse_refl / kPi * ((1.0f - f_out * 0.5f) * (1.0f - f_in * 0.5f));
    f3 h = out + in;
    if (!is_zero(h)) {
        h = normalize(h);
        float c_in = dot(in, h);
        D4 rr = rough * (2.0f * c_in * c_in);
        if (same_hemi && same_geo) val = val + diffuse_refl / kPi * rr * (f_out + f_in + f_out * f_in * (rr - 1.0f));
        if (cos_i != 0.0f && cos_o != 0.0f) {
            float c_out = dot(out, h);
            if (same_hemi && same_geo && c_in * c_out > 0.0f) {
                float c_tir = dot(in, h.z < 0.0f ? -h : h);
                float diel = fresnel_dielectric(c_tir, 1.0f, m.ior);
                float w = schlick_weight(c_tir);
                D3 schl = d3lerp(w, spec_r0, gray(1.0f));
                D3 fres = d3lerp(m.metallic, gray(diel), schl);
                val = val + spec_tint * (ggx.ndf(h) * ggx.g(out, in)) * fres / (4.0f * cos_i * cos_o);
            }
        }
    }
    float eta = out.z > 0.0f ? m.ior : srcp(m.ior);
    f3 ht = out + in * eta;
    if (cos_o != 0.0f && cos_i != 0.0f && !is_zero(ht) && m.spec_trans > 0.0f) {
        ht = normalize(ht);
        ht = ht.z < 0.0f ? -ht : ht;
        float c_out = dot(out, ht), c_in = dot(in, ht);
        float sqrt_denom = c_out + eta * c_in;
        f3 wh = !same_hemisphere(out, ht) ? -ht : ht;
        if (!same_hemi && c_out * c_in < 0.0f) {
            float F = fresnel_dielectric(c_out, 1.0f, m.ior);
            D4 num = ggx.ndf(wh) * ggx.g(out, in) *
                     (eta * eta * fmaxf(0.0f, dot(in, -wh)) * fmaxf(0.0f, dot(out, wh)));
            float denom = in.z * out.z * sqrt_denom * sqrt_denom;
            D4 q = num / denom;
            if (q.v < 0.0f) q = q * -1.0f;
            val = val + spec_trans * (1.0f - F) * q;
        }
    }
    return val;
}

// d log f / d p for the 3 colour channels and 4 parameters (0 where f = 0).
__device__ __forceinline__ void dlog_of(const D3& f, float* out12) {
    const D4* c[3] = {&f.x, &f.y, &f.z};
#pragma unroll
    for (int ch = 0; ch < 3; ++ch) {
        float inv = c[ch]->v != 0.0f ? srcp(c[ch]->v) : 0.0f;
#pragma unroll
        for (int k = 0; k < 4; ++k) out12[ch * 4 + k] = c[ch]->d[k] * inv;
    }
}

}  // namespace sspt
