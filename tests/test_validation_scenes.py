"""The reference's validation scenes (seesharp_amd.validation, after
SeeSharp.Validation) on the CPU: each builds a valid scene with the reference's
geometry, and the oracle's images of two independent seeds pass the
validator's 1 % mean bar (SeeSharp.Validation/Validator.cs:10-30)."""
import math

import numpy as np
import pytest

import oracle_lib as ol
from seesharp_amd import GlossyEmitter, validation
from test_oracle import lambert_polygon_irradiance


def test_scene_list_mirrors_the_validator():
    names = [v.Name for v in validation.SCENES if v.enabled]
    assert names == ["DirectIllumTransmit", "DirectIllumination", "SingleBounce", "MultiLight", "GlossyLight",
                     "Environment", "CornellBox", "TextureTest"]
    assert (validation.by_name("CornellBox").SamplesPerPixel, validation.by_name("CornellBox").MaxDepth) == (8, 5)
    counts = {v.Name: (v.MakeScene(32, 32).NumTriangles, len(v.MakeScene(32, 32).Emitters)) for v in validation.SCENES}
    assert counts == {"DirectIllumTransmit": (4, 2), "DirectIllumination": (4, 2), "SingleBounce": (6, 2),
                      "SingleBounceGlossy": (6, 2), "MultiLight": (6, 4), "GlossyLight": (4, 2),
                      "Environment": (2, 0), "CornellBox": (36, 2), "TextureTest": (6, 2)}
    gl = validation.by_name("GlossyLight").MakeScene(16, 16)
    assert all(isinstance(e, GlossyEmitter) and e.Exponent == 200.0 for e in gl.Emitters)


def test_means_agree_is_relative_to_the_first_image():
    a = np.ones((4, 4, 3))
    assert validation.means_agree([a, a * 1.009, a * 0.991])
    assert not validation.means_agree([a, a * 1.02])


@pytest.mark.parametrize("name", [v.Name for v in validation.SCENES if v.enabled])
def test_oracle_independent_seeds_pass_the_validator(name):
    v = validation.by_name(name)
    s = v.MakeScene(96, 96)
    assert s.IsValid
    o = ol.OracleScene(s)
    imgs = []
    for seed in (0xC030114, 571298512):
        img, _, nans = o.render(ol.settings(base_seed=seed, total_spp=4 * v.SamplesPerPixel, max_depth=v.MaxDepth),
                                0, 4 * v.SamplesPerPixel)
        assert nans == 0
        imgs.append(img)
    o.close()
    assert validation.image_mean(imgs[0]) > 0
    assert validation.means_agree(imgs, 0.02)


def lambert_irradiance_many(p, n, poly):
    """Lambert's polygon formula (unit radiance) at many points p (N x 3) with
    normal n: E = 1/2 sum_i beta_i (gamma_i . n), vectorised over the points."""
    e = np.zeros(p.shape[0])
    k = len(poly)
    for i in range(k):
        a = poly[i] - p
        b = poly[(i + 1) % k] - p
        a /= np.linalg.norm(a, axis=1, keepdims=True)
        b /= np.linalg.norm(b, axis=1, keepdims=True)
        beta = np.arccos(np.clip(np.sum(a * b, axis=1), -1, 1))
        g = np.cross(a, b)
        g /= np.linalg.norm(g, axis=1, keepdims=True)
        e += beta * (g @ n)
    return np.abs(0.5 * e)


def test_lambert_vectorised_matches_scalar():
    rng = np.random.default_rng(0)
    poly = np.array([[-0.1, -0.1, -1], [0.1, -0.1, -1], [0.1, 0.1, -1], [-0.1, 0.1, -1]], np.float64)
    p = np.c_[rng.uniform(-2, 2, (20, 2)), np.full(20, -2.0)]
    n = np.array([0, 0, 1.0])
    ref = [lambert_polygon_irradiance(q, n, poly) for q in p]
    assert np.allclose(lambert_irradiance_many(p, n, poly), ref, rtol=1e-12)


def test_oracle_multi_light_closed_form():
    """MultiLight: an albedo-1 diffuse plane at z = -2 under two 0.2 x 0.2
    lights of radiance 1000 at z = -1 facing it; at MaxDepth 2 every ground
    pixel is (1/pi) * 1000 * (sum of both lights' Lambert irradiances),
    evaluated at each iteration's jittered ground point and averaged."""
    s = validation.by_name("MultiLight").MakeScene(16, 16)
    polys = [m.Vertices.astype(np.float64) for m in s.Meshes[1:]]
    o = ol.OracleScene(s)
    spp = 256
    img, _, _ = o.render(ol.settings(total_spp=spp, max_depth=2), 0, spp)
    acc = np.zeros(256)
    ground = np.ones(256, bool)
    for it in range(spp):
        rays, hits = o.primary_hits(0xC030114, it)
        ground &= np.isin(hits["prim"], (0, 1))  # the ground's two triangles come first
        p = rays[:, :3].astype(np.float64) + hits["t"][:, None] * rays[:, 4:7].astype(np.float64)
        acc += sum(lambert_irradiance_many(p, np.array([0, 0, 1.0]), poly) for poly in polys)
    o.close()
    expected = 1000.0 * acc / spp / math.pi
    got = img.reshape(-1, 3).mean(axis=1)
    assert ground.sum() > 200
    ratio = got[ground] / expected[ground]
    # 256 samples of a four-triangle next-event choice: ~5 % noise per pixel
    assert np.all(np.abs(ratio - 1) < 0.2), ratio.min()
    assert abs(ratio.mean() - 1) < 0.01, ratio.mean()
