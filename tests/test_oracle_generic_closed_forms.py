"""CPU pins of the oracle's GenericMaterial value terms and of RNG.Next's output.

GenericMaterial.ComputeValueAndPdf (SeeSharp/Shading/Materials/GenericMaterial.cs:152-338)
with ComputeLocalParams (:354-381), TrowbridgeReitzDistribution (SeeSharp/Shading/
TrowbridgeReitzDistribution.cs:22-50) and Fresnel (SeeSharp/Shading/Fresnel.cs:4-48) is
checked three ways, none of which calls the oracle's own sampler or pdf:

* closed-form directional albedos. At IOR 1 the dielectric Fresnel term is 0, so a
  metallic-0, non-transmissive material is its diffuse and retro-reflection lobes only.
  At normal view (cos = 1, Schlick weight 0):
    diffuse   rho/pi * (1 - (1-mu)^5 / 2),  integral of that * mu over the hemisphere
              = rho * (1 - 1/42)                (int_0^1 (1-mu)^5 mu dmu = B(2,6) = 1/42);
    retro     rho/pi * Rr * F_i with Rr = 2 r cos^2(theta_d) = r (1 + mu),
              integral = 2 rho r (1/42 + 1/168) = rho * 5 r / 84;
  so the albedo is rho * (41/42 + 5 r / 84). At roughness 0 the retro lobe vanishes and
  the albedo at view cosine c is rho * (1 - (1-c)^5 / 2) * 41/42. A white metal (Schlick
  R0 = 1, so F = 1) at low roughness loses almost nothing to masking: albedo in [0.99, 1];
* point-by-point equality with the published formulas, written here from the papers
  in their own form: Burley 2012/2015's diffuse and retro-reflection lobes, the GGX
  distribution D = 1 / (pi a^2 cos^4 (1 + tan^2 / a^2)^2) (Walter et al. 2007) with the
  height-correlated Smith term G2 = 1 / (1 + L(o) + L(i)), L = (-1 + sqrt(1 + a^2 tan^2)) / 2
  (Heitz 2014, eqs. 72 and 99), Schlick's and the unpolarised dielectric Fresnel
  reflectance, and Walter et al.'s microfacet transmission (eq. 21) scaled by
  (eta_i / eta_o)^2, the camera-path convention of the reference (its `factor` is 1 off
  light subpaths);
* energy: a white dielectric that transmits everything it does not reflect has an
  albedo (reflection + transmission) close to 1 from either side.

The integrals use samplers written in this file (cosine hemisphere, GGX normals drawn
proportional to D * cos) with their own textbook densities.
"""
import ctypes as C
import math
import os
import subprocess

import numpy as np
import pytest

import oracle_lib as ol
from seesharp_amd import abi

L = ol.lib()
NORMAL = (0.0, 1.0, 0.0)


def generic(color, rough, metallic=0.0, trans=0.0, ior=1.5, tint=0.0):
    m = abi.MaterialDesc()
    m.kind = abi.MATERIAL_GENERIC
    m.base_color[:] = color
    m.base_color_texture = -1
    m.roughness = rough
    m.roughness_texture = -1
    m.metallic = metallic
    m.specular_tint_strength = tint
    m.specular_transmittance = trans
    m.index_of_refraction = ior
    m.anisotropic = 0.0
    return m


def eval_many(m, out_dir, dirs):
    """The oracle's f * |cos(in)| for n incident directions (n x 3, world, normal +y)."""
    dirs = np.ascontiguousarray(dirs, np.float32)
    n = dirs.shape[0]
    vc = np.zeros((n, 3), np.float32)
    o = (C.c_float * 3)(*out_dir)
    nrm = (C.c_float * 3)(*NORMAL)
    uv = (C.c_float * 2)(0.0, 0.0)
    L.or_material_eval_many(C.byref(m), None, nrm, uv, o, n, ol.fp(dirs), ol.fp(vc), None)
    return vc.astype(np.float64)


def view(theta_deg):
    t = math.radians(theta_deg)
    return np.array([math.sin(t), math.cos(t), 0.0])


# ---- samplers with their own densities (world frame, normal = +y) ----------------

def cos_hemisphere(rng, n):
    u1, u2 = rng.random(n), rng.random(n)
    r = np.sqrt(u1)
    phi = 2 * np.pi * u2
    d = np.stack([r * np.cos(phi), np.sqrt(np.maximum(0.0, 1 - u1)), r * np.sin(phi)], axis=1)
    return d, d[:, 1] / np.pi


def ggx_normals(rng, n, a):
    """Microfacet normals with density D(h) * cos(theta_h) (Walter et al. 2007, eqs. 35-36)."""
    u1, u2 = rng.random(n), rng.random(n)
    tan2 = a * a * u1 / (1.0 - u1)
    cos_t = 1.0 / np.sqrt(1.0 + tan2)
    sin_t = np.sqrt(np.maximum(0.0, 1.0 - cos_t * cos_t))
    phi = 2 * np.pi * u2
    h = np.stack([sin_t * np.cos(phi), cos_t, sin_t * np.sin(phi)], axis=1)
    return h, ggx_d(h, a) * cos_t


def reflect_samples(rng, n, a, o):
    h, pdf_h = ggx_normals(rng, n, a)
    oh = h @ o
    i = 2.0 * oh[:, None] * h - o[None, :]
    pdf = pdf_h / (4.0 * np.abs(oh))
    return i, pdf


# ---- the published formulas -----------------------------------------------------

def tan2(w):
    c2 = w[..., 1] ** 2
    return (1.0 - c2) / c2


def ggx_d(h, a):
    c = h[..., 1]
    t2 = tan2(h)
    return 1.0 / (np.pi * a * a * c ** 4 * (1.0 + t2 / (a * a)) ** 2)


def smith_lambda(w, a):
    return (-1.0 + np.sqrt(1.0 + a * a * tan2(w))) / 2.0


def smith_g2(o, i, a):
    return 1.0 / (1.0 + smith_lambda(o, a) + smith_lambda(i, a))


def schlick_weight(c):
    return np.clip(1.0 - c, 0.0, 1.0) ** 5


def fresnel_dielectric(cos_i, eta_i, eta_t):
    """Unpolarised Fresnel reflectance of a smooth dielectric boundary (textbook)."""
    cos_i = np.clip(cos_i, -1.0, 1.0)
    eta_i = np.broadcast_to(np.float64(eta_i), cos_i.shape).copy()
    eta_t = np.broadcast_to(np.float64(eta_t), cos_i.shape).copy()
    swap = cos_i < 0
    eta_i[swap], eta_t[swap] = eta_t[swap], eta_i[swap].copy()
    c = np.abs(cos_i)
    sin_t = eta_i / eta_t * np.sqrt(np.maximum(0.0, 1.0 - c * c))
    tir = sin_t >= 1.0
    cos_t = np.sqrt(np.maximum(0.0, 1.0 - sin_t * sin_t))
    rs = (eta_i * c - eta_t * cos_t) / (eta_i * c + eta_t * cos_t)
    rp = (eta_t * c - eta_i * cos_t) / (eta_t * c + eta_i * cos_t)
    return np.where(tir, 1.0, (rs * rs + rp * rp) / 2.0)


def luminance(c):
    # the build's RgbColor.Luminance rule (Rec. 709)
    return 0.2126 * c[0] + 0.7152 * c[1] + 0.0722 * c[2]


def generic_reflection(rho, r, metallic, tint, ior, o, i):
    """f * cos(in) of the reflection lobes, o and i on the normal's side (n x 3 each)."""
    rho = np.asarray(rho, np.float64)
    lum = luminance(rho)
    color_tint = rho / lum if lum > 0 else np.ones(3)
    spec_tint = (1 - tint) * np.ones(3) + tint * color_tint
    a = max(0.001, r * r)
    r0 = (1 - metallic) * ((ior - 1) / (ior + 1)) ** 2 * spec_tint + metallic * rho
    kd = rho * (1 - metallic)
    co, ci = o[..., 1], i[..., 1]
    fo, fi = schlick_weight(co), schlick_weight(ci)
    # Burley: diffuse (1 - F_L / 2)(1 - F_V / 2) and retro-reflection R_R (F_L + F_V + F_L F_V (R_R - 1))
    diffuse = kd[None, :] / np.pi * ((1 - fo / 2) * (1 - fi / 2))[:, None]
    h = o + i
    h /= np.linalg.norm(h, axis=-1, keepdims=True)
    cd = np.sum(i * h, axis=-1)
    rr = 2 * r * cd * cd
    retro = kd[None, :] / np.pi * (rr * (fo + fi + fo * fi * (rr - 1)))[:, None]
    # GGX reflection with Fresnel mixed between the dielectric and Schlick's
    f_diel = fresnel_dielectric(cd, 1.0, ior)[:, None]
    f_schlick = r0[None, :] + (1 - r0[None, :]) * schlick_weight(cd)[:, None]
    fres = (1 - metallic) * f_diel + metallic * f_schlick
    spec = spec_tint[None, :] * (ggx_d(h, a) * smith_g2(o, i, a) / (4 * ci * co))[:, None] * fres
    return (diffuse + retro + spec) * ci[:, None]


def generic_transmission(rho, r, trans, ior, o, i):
    """f * |cos(in)| of the transmission lobe, o and i on opposite sides (n x 3)."""
    a = max(0.001, r * r)
    outside = o[..., 1] > 0
    eta_o = np.where(outside, 1.0, ior)
    eta_i = np.where(outside, ior, 1.0)
    h = -(eta_o[:, None] * o + eta_i[:, None] * i)
    h /= np.linalg.norm(h, axis=-1, keepdims=True)
    h = np.where(h[:, 1:2] < 0, -h, h)  # D and Lambda are even in h
    oh, ih = np.sum(o * h, -1), np.sum(i * h, -1)
    co, ci = o[..., 1], i[..., 1]
    fres = fresnel_dielectric(oh, 1.0, ior)
    walter = (np.abs(ih) * np.abs(oh) / (np.abs(ci) * np.abs(co)) * eta_o ** 2 * (1 - fres)
              * smith_g2(o, i, a) * ggx_d(h, a) / (eta_i * ih + eta_o * oh) ** 2)
    val = walter * (eta_i / eta_o) ** 2 * trans
    # Walter's chi+ terms: each direction lies on the side of the microfacet that its
    # own side of the macro surface sees
    val = np.where((oh / co > 0) & (ih / ci > 0), val, 0.0)
    return val[:, None] * np.asarray(rho, np.float64)[None, :] * np.abs(ci)[:, None]


# ---- closed-form albedos ----------------------------------------------------------

RHO = np.array([0.9, 0.5, 0.2])


@pytest.mark.parametrize("rough", [0.0, 0.3, 0.7, 1.0])
def test_normal_view_albedo_ior1_is_diffuse_plus_retro_closed_form(rough):
    m = generic(RHO, rough, ior=1.0)
    rng = np.random.default_rng(11)
    d, pdf = cos_hemisphere(rng, 2_000_000)
    est = eval_many(m, (0.0, 1.0, 0.0), d) / pdf[:, None]
    mean, se = est.mean(0), est.std(0) / math.sqrt(est.shape[0])
    expect = RHO * (41 / 42 + 5 * rough / 84)
    assert np.all(np.abs(mean - expect) <= 5 * se + 2e-5), (mean, expect, se)


@pytest.mark.parametrize("theta", [30.0, 60.0, 80.0])
def test_oblique_view_albedo_ior1_roughness0_closed_form(theta):
    m = generic(RHO, 0.0, ior=1.0)
    o = view(theta)
    rng = np.random.default_rng(12)
    d, pdf = cos_hemisphere(rng, 1_000_000)
    est = eval_many(m, tuple(o), d) / pdf[:, None]
    mean, se = est.mean(0), est.std(0) / math.sqrt(est.shape[0])
    expect = RHO * (1 - 0.5 * (1 - o[1]) ** 5) * 41 / 42
    assert np.all(np.abs(mean - expect) <= 5 * se + 2e-5), (mean, expect, se)


@pytest.mark.parametrize("rough", [0.05, 0.1, 0.2])
@pytest.mark.parametrize("theta", [0.0, 45.0, 70.0])
def test_white_metal_albedo_is_nearly_one(theta, rough):
    """White metal (F = 1): the albedo is the Smith masking's single-scattering share,
    within 1 % of 1 at these roughnesses, and equals the published formulas' albedo
    on the same samples. At roughness 0.05 (alpha 0.0025) the reference's fp32
    sin^2 = 1 - cos^2 (SeeSharp/Shading/ShadingSpace.cs:83) rounds away most of
    tan^2 of the sampled normals; D is convex in tan^2, so its albedo lands about
    0.1 % above 1 there: the bound allows 0.2 %."""
    m = generic((1.0, 1.0, 1.0), rough, metallic=1.0)
    o = view(theta)
    rng = np.random.default_rng(13)
    i, pdf = reflect_samples(rng, 400_000, rough ** 2, o)
    ok = i[:, 1] > 0
    est = np.zeros((i.shape[0], 3))
    est[ok] = eval_many(m, tuple(o), i[ok]) / pdf[ok, None]
    mean, se = est.mean(0), est.std(0) / math.sqrt(est.shape[0])
    upper = 1.002 if rough < 0.1 else 1.0 + 5 * se
    assert np.all(mean <= upper) and np.all(mean >= 0.99), (mean, se)
    if rough >= 0.1:
        want = np.zeros((i.shape[0], 3))
        want[ok] = generic_reflection((1, 1, 1), rough, 1.0, 0.0, 1.5, np.broadcast_to(o, i[ok].shape),
                                      i[ok]) / pdf[ok, None]
        assert np.all(np.abs(mean - want.mean(0)) <= 1e-4), (mean, want.mean(0))


# ---- point-by-point against the published formulas ------------------------------------

def random_upper(rng, n):
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    d[:, 1] = np.abs(d[:, 1])
    return d[d[:, 1] > 0.05]


@pytest.mark.parametrize("params", [
    dict(rho=(0.9, 0.5, 0.2), r=0.5, metallic=0.0, tint=0.0, ior=1.5),
    dict(rho=(0.9, 0.5, 0.2), r=0.3, metallic=0.0, tint=1.0, ior=1.5),
    dict(rho=(0.2, 0.7, 0.4), r=0.8, metallic=1.0, tint=0.0, ior=1.5),
    dict(rho=(0.6, 0.6, 0.3), r=0.4, metallic=0.5, tint=0.0, ior=1.33),
    dict(rho=(0.3, 0.3, 0.9), r=1.0, metallic=0.0, tint=0.0, ior=2.0),
])
def test_reflection_value_equals_published_formulas(params):
    m = generic(params["rho"], params["r"], metallic=params["metallic"], tint=params["tint"], ior=params["ior"])
    rng = np.random.default_rng(14)
    for _ in range(4):
        o = random_upper(rng, 16)[0]
        i = random_upper(rng, 4000)
        got = eval_many(m, tuple(o.astype(np.float32)), i)
        o32 = np.broadcast_to(o.astype(np.float32).astype(np.float64), i.shape)
        i32 = i.astype(np.float32).astype(np.float64)
        want = generic_reflection(params["rho"], params["r"], params["metallic"], params["tint"], params["ior"],
                                  o32, i32)
        assert np.allclose(got, want, rtol=3e-4, atol=1e-6), np.max(np.abs(got - want) / (np.abs(want) + 1e-6))


@pytest.mark.parametrize("params", [
    dict(rho=(1.0, 1.0, 1.0), r=0.5, trans=1.0, ior=1.5),
    dict(rho=(0.8, 0.4, 0.9), r=0.3, trans=0.9, ior=1.33),
])
@pytest.mark.parametrize("side", [1.0, -1.0])
def test_transmission_value_equals_published_formula(params, side):
    m = generic(params["rho"], params["r"], trans=params["trans"], ior=params["ior"])
    rng = np.random.default_rng(15)
    for _ in range(4):
        o = random_upper(rng, 16)[0] * np.array([1.0, side, 1.0])
        i = random_upper(rng, 4000) * np.array([1.0, -side, 1.0])
        got = eval_many(m, tuple(o.astype(np.float32)), i)
        o32 = np.broadcast_to(o.astype(np.float32).astype(np.float64), i.shape)
        i32 = i.astype(np.float32).astype(np.float64)
        want = generic_transmission(params["rho"], params["r"], params["trans"], params["ior"], o32, i32)
        assert np.allclose(got, want, rtol=3e-4, atol=1e-6), np.max(np.abs(got - want) / (np.abs(want) + 1e-6))


@pytest.mark.parametrize("side", [1.0, -1.0])
def test_white_dielectric_reflects_or_transmits_nearly_everything(side):
    """Reflection + transmission of a white, fully transmissive dielectric (camera-path
    convention): its albedo stays within the masking loss of 1 on both sides."""
    m = generic((1.0, 1.0, 1.0), 0.5, trans=1.0, ior=1.5)
    o = view(20.0) * np.array([1.0, side, 1.0])
    rng = np.random.default_rng(16)
    n = 2_000_000
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    est = eval_many(m, tuple(o), d).mean(1) * (4 * np.pi)
    mean, se = est.mean(), est.std() / math.sqrt(n)
    assert 0.9 <= mean <= 1.0 + 5 * se, (mean, se)


# ---- RNG.Next: full-domain check of the output permutation --------------------------

RNG_CHECK = r"""
#include <pthread.h>
#include <stdint.h>
#include <stdio.h>
void or_rng_stream(uint32_t seed, uint32_t n, uint32_t* out_u32, float* out_float);
/* PCG's RXS-M-XS 32/32 output (O'Neill 2014, pcg_output_rxs_m_xs_32_32), inverted step
 * by step: x ^= x >> 22, x *= 277803737, x ^= x >> ((x >> 28) + 4). */
static uint32_t inv_mult;
static uint32_t unxorshift(uint32_t y, unsigned k) {
    uint32_t x = y;
    for (unsigned s = 0; s < 32; s += k) x = y ^ (x >> k);
    return x;
}
static uint32_t invert(uint32_t out) {
    uint32_t word = out ^ (out >> 22);
    uint32_t t = word * inv_mult;
    return unxorshift(t, (t >> 28) + 4);
}
static unsigned long long bad[8];
static void* run(void* arg) {
    uintptr_t id = (uintptr_t)arg;
    uint64_t lo = (uint64_t)id << 29, hi = lo + (1ull << 29);
    for (uint64_t s = lo; s < hi; ++s) {
        uint32_t u;
        or_rng_stream((uint32_t)s, 1, &u, 0);
        if (invert(u) != (uint32_t)s) ++bad[id];
    }
    return 0;
}
int main(void) {
    uint32_t a = 277803737u, x = a;
    for (int k = 0; k < 5; ++k) x *= 2u - a * x;
    if (a * x != 1u) return 2;
    inv_mult = x;
    pthread_t th[8];
    for (uintptr_t t = 0; t < 8; ++t) pthread_create(&th[t], 0, run, (void*)t);
    unsigned long long total = 0;
    for (int t = 0; t < 8; ++t) { pthread_join(th[t], 0); total += bad[t]; }
    printf("%llu\n", total);
    return total != 0;
}
"""


def test_rng_next_output_is_the_pcg_rxs_m_xs_permutation_on_all_states(tmp_path):
    """RNG.Next (SeeSharp/Sampling/RNG.cs:90-94) returns the RXS-M-XS output of the
    state before the LCG step. Inverting that published output function maps the first
    Next() of every one of the 2^32 seeds back to its seed, so the oracle's output is
    exactly that bijection (the reference's files hold no output vector of their own)."""
    src = tmp_path / "rng_check.c"
    src.write_text(RNG_CHECK)
    exe = tmp_path / "rng_check"
    subprocess.run(["gcc", "-O2", "-pthread", "-o", str(exe), str(src), "-L" + ol.ORACLE_DIR, "-loracle",
                    "-Wl,-rpath," + ol.ORACLE_DIR], check=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=600)
    assert r.returncode == 0 and r.stdout.strip() == "0", (r.returncode, r.stdout, r.stderr)
