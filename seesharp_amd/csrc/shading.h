// shading.h -- device functions of the surface interaction, the materials and
// the environment map.
//
//   surface()            TinyEmbree Hit -> SurfacePoint (SeeSharp/Geometry/SurfacePoint.cs,
//                        Mesh.ComputeErrorOffset SeeSharp/Geometry/Mesh.cs:50-60)
//   generic_*            GenericMaterial (SeeSharp/Shading/Materials/GenericMaterial.cs)
//   diffuse_*            DiffuseMaterial (SeeSharp/Shading/Materials/DiffuseMaterial.cs)
//   env_*                EnvironmentMap (SeeSharp/Shading/Background/EnvironmentMap.cs) over
//                        RegularGrid2d / PiecewiseConstantPDF (SeeSharp/Sampling/)
#pragma once

#include "device_math.h"
#include "scene_types.h"

namespace sspt {

struct Surf {
    f3 pos, ng, ns;
    f2 uv;
    float err;
    int mat;
};

// pos: the triangle's 3 tri_pos entries, aux: its 4 tri_aux entries (HBM or LDS)
__device__ __forceinline__ Surf surface_at(const float4* pos, const float4* aux, float u, float v) {
    Surf s;
    float4 a = pos[0], b = pos[1], c = pos[2];
    f3 p0 = f3_of(a), p1 = f3_of(b), p2 = f3_of(c);
    float w0 = 1.0f - u - v;
    s.pos = u * p1 + v * p2 + w0 * p0;
    s.ng = normalize(cross(p1 - p0, p2 - p0));
    float4 x0 = aux[0], x1 = aux[1], x2 = aux[2], x3 = aux[3];
    int info = __float_as_int(x3.w);
    s.mat = info & kInfoMaterialMask;
    s.ns = (info & kInfoHasNormals) ? normalize(u * f3_of(x1) + v * f3_of(x2) + w0 * f3_of(x0)) : s.ng;
    if (info & kInfoHasUvs) {
        // uv0 = (x0.w, x1.w), uv1 = (x2.w, x3.x), uv2 = (x3.y, x3.z)
        s.uv = {u * x2.w + v * x3.y + w0 * x0.w, u * x3.x + v * x3.z + w0 * x1.w};
    } else {
        s.uv = {0.0f, 0.0f};
    }
    f3 diag = vabs(u * p1) + vabs(v * p2) + vabs(w0 * p0);
    s.err = length(diag) * 32.0f * 1.19209e-07f;
    return s;
}
__device__ __forceinline__ Surf surface(const DevScene& S, int li, float u, float v) {
    return surface_at(S.tri_pos + 3 * li, S.tri_aux + 4 * li, u, v);
}

// ImageTexture.ComputeTexel (SeeSharp/Images/ImageTexture.cs:27-44): nearest texel.
__device__ __forceinline__ int texel_index(const DevTexture& t, f2 uv) {
    int col = to_int(uv.x * (float)t.w);
    int row = to_int(uv.y * (float)t.h);
    if (t.border == 0) {
        row = (row % t.h + t.h) % t.h;
        col = (col % t.w + t.w) % t.w;
    } else {
        row = min(max(row, 0), t.h - 1);
        col = min(max(col, 0), t.w - 1);
    }
    return row * t.w + col;
}
__device__ __forceinline__ f3 lookup_rgb(const DevScene& S, int tex, f3 constant, f2 uv) {
    if (tex < 0) return constant;
    DevTexture t = S.textures[tex];
    const float* p = S.texels + t.offset + (uint64_t)texel_index(t, uv) * 3;
    return {p[0], p[1], p[2]};
}
__device__ __forceinline__ float lookup_mono(const DevScene& S, int tex, float constant, f2 uv) {
    if (tex < 0) return constant;
    DevTexture t = S.textures[tex];
    return S.texels[t.offset + (uint64_t)texel_index(t, uv)];
}

// Shading context of one hit (SurfaceShader / ShadingContext).
struct Shader {
    Frame f;
    f3 out_local, out_world, ng;
    DevMaterial m;
    // GenericMaterial.LocalParams (GenericMaterial.cs:340-381); base color for diffuse.
    f3 base, spec_tint, diffuse_refl, spec_trans, spec_r0;
    float roughness, ax, ay;
    bool light;  // ShadingContext.IsOnLightSubpath
};

__device__ __forceinline__ void make_shader(const DevScene& S, const Surf& s, f3 out_world, Shader* sh) {
    sh->m = S.materials[s.mat];
    sh->light = false;
    sh->f.build(s.ns);
    sh->out_world = out_world;
    sh->out_local = sh->f.to_local(out_world);
    sh->ng = s.ng;
    const DevMaterial& m = sh->m;
    f3 base = lookup_rgb(S, m.base_tex, mk3(m.base[0], m.base[1], m.base[2]), s.uv);
    sh->base = base;
    if (m.kind == 1) {
        float lum = luminance(base);
        f3 color_tint = lum > 0.0f ? base / lum : gray(1.0f);
        sh->spec_tint = lerp(m.spec_tint, gray(1.0f), color_tint);
        sh->roughness = lookup_mono(S, m.rough_tex, m.roughness, s.uv);
        float aspect = ssqrt(1.0f - m.anisotropic * 0.9f);
        sh->ax = fmaxf(0.001f, sdiv(sh->roughness * sh->roughness, aspect));
        sh->ay = fmaxf(0.001f, sh->roughness * sh->roughness * aspect);
        sh->spec_r0 = lerp(m.metallic, schlick_r0(m.ior) * sh->spec_tint, base);
        float dw = (1.0f - m.metallic) * (1.0f - m.spec_trans);
        sh->diffuse_refl = base * dw;
        sh->spec_trans = m.spec_trans * base;
    }
}

// Material.ShouldReflect (SeeSharp/Shading/Materials/Material.cs:127-134)
__device__ __forceinline__ bool should_reflect(f3 ng, f3 out_dir, f3 in_dir) {
    return dot(out_dir, ng) * dot(in_dir, ng) >= 0.0f;
}

// GenericMaterial.ComputeValueAndPdf (GenericMaterial.cs:152-338), forward pdf only:
// the path tracer never uses the reverse pdf (PathTracer.cs:510-512).
__device__ void generic_value_pdf(const Shader& sh, f3 in, f3* value, float* pdf_out) {
    const DevMaterial& m = sh.m;
    Ggx ggx{sh.ax, sh.ay};
    f3 out = sh.out_local;
    float cos_o = fabsf(out.z), cos_i = fabsf(in.z);
    bool same_geo = should_reflect(sh.ng, sh.out_world, sh.f.to_world(in));
    bool same_hemi = same_hemisphere(out, in);
    f3 val = gray(0.0f);
    float pdf0 = 0.0f, pdf1 = 0.0f, pdf2 = 0.0f;

    float f_out = schlick_weight(cos_o), f_in = schlick_weight(cos_i);
    if (same_hemi) {
        if (same_geo) val = val + sh.diffuse_refl / kPi * (1.0f - f_out * 0.5f) * (1.0f - f_in * 0.5f);
        pdf0 = cos_i * kInvPi;
    }
    f3 h = out + in;
    if (!is_zero(h)) {
        h = normalize(h);
        float c_in = dot(in, h);
        float rr = 2.0f * sh.roughness * c_in * c_in;
        if (same_hemi && same_geo)
            val = val + sh.diffuse_refl / kPi * rr * (f_out + f_in + f_out * f_in * (rr - 1.0f));
        if (cos_i != 0.0f && cos_o != 0.0f) {
            float c_out = dot(out, h);
            if (same_hemi && same_geo && c_in * c_out > 0.0f) {
                float c_tir = dot(in, h.z < 0.0f ? -h : h);
                f3 diel = gray(fresnel_dielectric(c_tir, 1.0f, m.ior));
                f3 schl = schlick_fresnel(sh.spec_r0, c_tir);
                f3 fres = lerp(m.metallic, diel, schl);
                val = val + sh.spec_tint * ggx.ndf(h) * ggx.g(out, in) * fres / (4.0f * cos_i * cos_o);
            }
            float jf = fabsf(4.0f * c_out), jr = fabsf(4.0f * c_in);
            if (jf != 0.0f && jr != 0.0f) pdf1 = sdiv(ggx.pdf(out, h), jf);
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
            float num = ggx.ndf(wh) * ggx.g(out, in);
            num *= eta * eta * fmaxf(0.0f, dot(in, -wh)) * fmaxf(0.0f, dot(out, wh));
            if (sh.light) {  // GenericMaterial.cs:255: importance, not radiance, is transported
                const float factor = srcp(eta);
                num *= factor * factor;
            }
            float denom = in.z * out.z * sqrt_denom * sqrt_denom;
            val = val + (gray(1.0f) - gray(F)) * sh.spec_trans * fabsf(sdiv(num, denom));
        }
        if (sqrt_denom != 0.0f) {
            float jac = sdiv(eta * eta * fmaxf(0.0f, dot(in, -wh)), sqrt_denom * sqrt_denom);
            pdf2 += ggx.pdf(out, wh) * jac;
        }
    }
    float schl = average(schlick_fresnel(sh.spec_r0, fabsf(out.z)));
    float bias = fmaxf(1.0f - schl, 0.75f);
    float dw = bias * (1.0f - m.metallic) * (1.0f - m.spec_trans);
    float fr = fresnel_dielectric(dot(out, h.z < 0.0f ? -h : h), 1.0f, m.ior);
    float sel_refl = (1.0f - dw) * (1.0f - (1.0f - fr) * m.spec_trans);
    float ft = fresnel_dielectric(dot(out, ht), 1.0f, m.ior);
    float sel_trans = (1.0f - dw) * (1.0f - ft) * m.spec_trans;
    *value = val;
    *pdf_out = pdf0 * dw + pdf1 * sel_refl + pdf2 * sel_trans;
}

// The reverse pdf of GenericMaterial.ComputeValueAndPdf (GenericMaterial.cs:152-338):
// the density of sampling `out` from `in` per lobe (pdfsRev), weighted with the
// selection probabilities of the forward direction (selectTransmitIn repeats
// (1 - fresnelT) of `out`, :311). Only the bidirectional MIS weights read it.
__device__ float generic_pdf_rev(const Shader& sh, f3 in) {
    const DevMaterial& m = sh.m;
    Ggx ggx{sh.ax, sh.ay};
    f3 out = sh.out_local;
    float cos_o = fabsf(out.z), cos_i = fabsf(in.z);
    float rev0 = 0.0f, rev1 = 0.0f, rev2 = 0.0f;
    if (same_hemisphere(out, in)) rev0 = cos_o * kInvPi;
    f3 h = out + in;
    if (!is_zero(h)) {
        h = normalize(h);
        if (cos_i != 0.0f && cos_o != 0.0f) {
            float jf = fabsf(4.0f * dot(out, h)), jr = fabsf(4.0f * dot(in, h));
            if (jf != 0.0f && jr != 0.0f) rev1 = sdiv(ggx.pdf(in, h), jr);
        }
    }
    float eta = out.z > 0.0f ? m.ior : srcp(m.ior);
    f3 ht = out + in * eta;
    if (cos_o != 0.0f && cos_i != 0.0f && !is_zero(ht) && m.spec_trans > 0.0f) {
        ht = normalize(ht);
        ht = ht.z < 0.0f ? -ht : ht;
        float eta_in = in.z > 0.0f ? m.ior : srcp(m.ior);
        f3 hin = normalize(out * eta_in + in);
        hin = hin.z < 0.0f ? -hin : hin;
        if (!is_zero(hin)) {
            f3 wh = !same_hemisphere(in, hin) ? -hin : hin;
            float sqrt_denom_in = dot(hin, in) + eta_in * dot(out, hin);
            if (sqrt_denom_in != 0.0f) {
                float jac = sdiv(eta_in * eta_in * fmaxf(0.0f, dot(out, -wh)), sqrt_denom_in * sqrt_denom_in);
                rev2 += ggx.pdf(in, wh) * jac;
            }
        }
    }
    float schl = average(schlick_fresnel(sh.spec_r0, fabsf(out.z)));
    float bias = fmaxf(1.0f - schl, 0.75f);
    float dw = bias * (1.0f - m.metallic) * (1.0f - m.spec_trans);
    float fr = fresnel_dielectric(dot(out, h.z < 0.0f ? -h : h), 1.0f, m.ior);
    float sel_refl = (1.0f - dw) * (1.0f - (1.0f - fr) * m.spec_trans);
    float ft = fresnel_dielectric(dot(out, ht), 1.0f, m.ior);
    float sel_trans = (1.0f - dw) * (1.0f - ft) * m.spec_trans;
    return rev0 * dw + rev1 * sel_refl + rev2 * sel_trans;
}

// DiffuseMaterial.Evaluate / Pdf (DiffuseMaterial.cs:127-139,167-188)
__device__ __forceinline__ void diffuse_value_pdf(const Shader& sh, f3 in_world, f3* value, float* pdf) {
    bool refl = should_reflect(sh.ng, sh.out_world, in_world);
    f3 in = sh.f.to_local(in_world);
    bool same = same_hemisphere(sh.out_local, in);
    f3 v = gray(0.0f);
    if (sh.m.transmitter && !refl) {
        if (!same) v = sh.base / kPi;
    } else if (refl) {
        if (same) v = sh.base / kPi;
    }
    *value = v;
    float pr = same ? fabsf(in.z) * kInvPi : 0.0f;
    if (sh.m.transmitter) {
        float pt = !same ? fabsf(in.z) * kInvPi : 0.0f;
        *pdf = (pr + pt) * 0.5f;
    } else {
        *pdf = pr;
    }
}

// DiffuseBsdf.Pdf / DiffuseTransmission.Pdf (DiffuseMaterial.cs:63-71,95-103): the
// reverse pdf, |cos| of the outgoing direction / pi on the lobe's side.
__device__ __forceinline__ float diffuse_pdf_rev(const Shader& sh, f3 in_world) {
    f3 in = sh.f.to_local(in_world);
    bool same = same_hemisphere(sh.out_local, in);
    float pr = same ? fabsf(sh.out_local.z) * kInvPi : 0.0f;
    if (sh.m.transmitter) {
        float pt = !same ? fabsf(sh.out_local.z) * kInvPi : 0.0f;
        return (pr + pt) * 0.5f;
    }
    return pr;
}

// SurfaceShader.Pdf(inDir) (SurfaceShader.cs:64-67): forward and reverse pdfs.
__device__ __forceinline__ void pdf_pair(const Shader& sh, f3 in_world, float* fwd, float* rev) {
    f3 v;
    if (sh.m.kind == 1) {
        f3 in = sh.f.to_local(in_world);
        generic_value_pdf(sh, in, &v, fwd);
        *rev = generic_pdf_rev(sh, in);
    } else {
        diffuse_value_pdf(sh, in_world, &v, fwd);
        *rev = diffuse_pdf_rev(sh, in_world);
    }
}

// SurfaceShader.EvaluateWithCosine + Pdf(..).Pdf for a world direction.
__device__ __forceinline__ void eval_cos_pdf(const Shader& sh, f3 in_world, f3* value_cos, float* pdf) {
    if (sh.m.kind == 1) {
        f3 in = sh.f.to_local(in_world);
        f3 v;
        generic_value_pdf(sh, in, &v, pdf);
        *value_cos = v * fabsf(in.z);
    } else {
        f3 v;
        diffuse_value_pdf(sh, in_world, &v, pdf);
        *value_cos = v * fabsf(sh.f.to_local(in_world).z);
    }
}

struct BsdfSample {
    f3 dir;
    float pdf;
    f3 weight;
    float pdf_rev;  // BsdfSample.PdfReverse: set by sample_bsdf<true> only
};

// SurfaceShader.Sample (GenericMaterial.cs:92-150, DiffuseMaterial.cs:141-165);
// REV also computes the reverse pdf of the sampled direction.
template <bool REV = false>
__device__ BsdfSample sample_bsdf(const Shader& sh, float pc, f2 pd) {
    BsdfSample bad{gray(0.0f), 0.0f, gray(0.0f), 0.0f};
    const DevMaterial& m = sh.m;
    if (m.kind == 1) {
        f3 out = sh.out_local;
        if (out.z == 0.0f) return bad;
        Ggx ggx{sh.ax, sh.ay};
        float eta = out.z > 0.0f ? srcp(m.ior) : m.ior;
        float schl = average(schlick_fresnel(sh.spec_r0, fabsf(out.z)));
        float bias = fmaxf(1.0f - schl, 0.75f);
        float dw = bias * (1.0f - m.metallic) * (1.0f - m.spec_trans);
        f3 in;
        if (pc <= dw) {
            in = to_cos_hemisphere(pd);
            if (out.z < 0.0f) in.z *= -1.0f;
        } else {
            f3 h = ggx.sample(out, pd);
            float c_out = dot(out, h.z < 0.0f ? -h : h);
            float fres = fresnel_dielectric(c_out, 1.0f, m.ior);
            float sel_t = (1.0f - fres) * m.spec_trans;
            if (sdiv(pc - dw, 1.0f - dw) < sel_t) {
                if (dot(out, h) < 0.0f) return bad;
                f3 r;
                if (!refract(out, h, eta, &r)) return bad;
                in = r;
            } else {
                in = reflect(out, h);
            }
        }
        f3 val;
        float pdf;
        generic_value_pdf(sh, in, &val, &pdf);
        BsdfSample bs;
        bs.pdf = pdf;
        bs.weight = val * fabsf(in.z) / pdf;
        bs.dir = sh.f.to_world(in);
        bs.pdf_rev = REV ? generic_pdf_rev(sh, in) : 0.0f;
        return bs;
    }
    f3 d = to_cos_hemisphere(pd);
    if (m.transmitter && pc < 0.5f) {
        if (sh.out_local.z > 0.0f) d.z *= -1.0f;
    } else {
        if (sh.out_local.z < 0.0f) d.z *= -1.0f;
    }
    f3 dir = sh.f.to_world(d);
    f3 vc;
    float pdf;
    eval_cos_pdf(sh, dir, &vc, &pdf);
    if (pdf == 0.0f) return bad;
    return {dir, pdf, vc / pdf, REV ? diffuse_pdf_rev(sh, dir) : 0.0f};
}

// ------------------------------------------------------------ environment --
// The scene's environment map: texels and the RegularGrid2d CDFs live in the
// world's pools at the offsets of its SceneRec.
struct Env {
    const float4* rgb;
    const float* row_cdf;
    const float* col_cdf;
    const int* row_guide;
    const int* col_guide;
    int w, h;
    int row_g, col_g;
};
__device__ __forceinline__ Env env_of(const DevScene& S, const SceneRec& R) {
    return Env{S.env_rgb + R.env_rgb_off,
               S.env_cdf + R.env_row_off,
               S.env_cdf + R.env_col_off,
               S.env_guide + R.env_guide_row_off,
               S.env_guide + R.env_guide_col_off,
               R.env_w,
               R.env_h,
               R.env_guide_row_g,
               R.env_guide_col_g};
}
__device__ __forceinline__ f2 world_to_spherical(f3 dir) {
    dir = normalize(dir);
    f2 sp{atan2f(dir.z, dir.x), atan2f(ssqrt(dir.x * dir.x + dir.z * dir.z), dir.y)};
    if (sp.x < 0.0f) sp.x += kPi * 2.0f;
    return sp;
}
__device__ __forceinline__ f3 env_pixel(const Env& E, int col, int row) {
    col = min(max(col, 0), E.w - 1);
    row = min(max(row, 0), E.h - 1);
    return f3_of(E.rgb[row * E.w + col]);
}
__device__ __forceinline__ f3 env_radiance(const Env& E, f3 dir) {
    f2 sp = world_to_spherical(dir);
    float px = sp.x * kInv2Pi, py = sp.y * kInvPi;
    return env_pixel(E, to_int(px * (float)E.w), to_int(py * (float)E.h));
}
__device__ __forceinline__ float cdf_prob(const float* cdf, int i) {
    return i > 0 ? cdf[i] - cdf[i - 1] : cdf[i];
}
// RegularGrid2d.Pdf (SeeSharp/Sampling/RegularGrid2d.cs:39-47)
__device__ __forceinline__ float grid_pdf(const Env& E, float x, float y) {
    int row = min(max(to_int(y * (float)E.h), 0), E.h - 1);
    int col = min(max(to_int(x * (float)E.w), 0), E.w - 1);
    float p = cdf_prob(E.row_cdf, row);
    if (p == 0.0f) return 0.0f;
    p *= cdf_prob(E.col_cdf + (size_t)row * E.w, col);
    return p * (float)(E.h * E.w);
}
__device__ __forceinline__ float env_pdf(const Env& E, f3 dir) {
    f2 sp = world_to_spherical(dir);
    return sdiv(grid_pdf(E, sp.x * kInv2Pi, sp.y * kInvPi), ssin(sp.y) * kPi * kPi * 2.0f);
}
// PiecewiseConstantPDF.Sample (SeeSharp/Sampling/PiecewiseConstant.cs:39-51): the
// lower bound of u in the CDF (first entry >= u). The guide table holds the
// lower bounds of j / g (g a power of two, so j = floor(u * g) is exact and
// j / g <= u <= (j + 1) / g); the lower bound of u lies between guide[j] and
// guide[j + 1] and is searched there only: the same index as a search of the
// whole CDF, with a few dependent loads instead of log2(n).
__device__ __forceinline__ int lower_bound(const float* cdf, const int* guide, int g, float u) {
    const int j = min((int)(u * (float)g), g - 1);
    int lo = guide[j], len = guide[j + 1] - lo;
    while (len > 0) {
        int half = len >> 1;
        if (cdf[lo + half] < u) {
            lo += half + 1;
            len -= half + 1;
        } else {
            len = half;
        }
    }
    return lo;
}
struct EnvSample {
    f3 dir;
    float pdf;
    f3 weight;
};
// EnvironmentMap.SampleDirection (EnvironmentMap.cs:40-73)
__device__ EnvSample env_sample(const Env& E, f2 u) {
    int ri = min(lower_bound(E.row_cdf, E.row_guide, E.row_g, u.y), E.h - 1);
    float rlo = ri == 0 ? 0.0f : E.row_cdf[ri - 1];
    float rrel = sdiv(u.y - rlo, E.row_cdf[ri] - rlo);
    const float* ccdf = E.col_cdf + (size_t)ri * E.w;
    const int* cguide = E.col_guide + (size_t)ri * (E.col_g + 1);
    int ci = min(lower_bound(ccdf, cguide, E.col_g, u.x), E.w - 1);
    float clo = ci == 0 ? 0.0f : ccdf[ci - 1];
    float crel = sdiv(u.x - clo, ccdf[ci] - clo);
    float px = sdiv((float)ci + crel, (float)E.w);
    float py = sdiv((float)ri + rrel, (float)E.h);
    float pdf = grid_pdf(E, px, py);
    float phi = px * 2.0f * kPi, theta = py * kPi;
    float sin_t, cos_t, sin_p, cos_p;
    ssincos(theta, &sin_t, &cos_t);
    ssincos(phi, &sin_p, &cos_p);
    EnvSample r;
    r.dir = mk3(sin_t * cos_p, cos_t, sin_t * sin_p);
    float jac = sin_t * kPi * kPi * 2.0f;
    if (jac == 0.0f) {
        r.pdf = 0.0f;
        r.weight = gray(0.0f);
        return r;
    }
    pdf /= jac;
    r.pdf = pdf;
    r.weight = env_pixel(E, to_int(px * (float)E.w), to_int(py * (float)E.h)) / pdf;
    return r;
}

}  // namespace sspt
