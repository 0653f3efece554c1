"""The oracle's light tracer (VertexCacheBidir as a pure light tracer:
BidirBase.Light.cs:35-277, RandomWalk.cs, DiffuseEmitter.cs:44-64,
GlossyEmitter.cs:47-70, PerspectiveCamera.SampleResponse) pinned against the
oracle's path tracer, an independent estimator of the same image.

A light tracer splats every light-path vertex onto the camera, so it estimates
the image of all paths with at least two edges; the path tracer with MinDepth 2
estimates the same. Both are unbiased, so their images agree within their
Monte Carlo error (z-scores from independent seed batches). The scenes cover
what the light side does differently from the camera side: the emitter's ray
sampling (diffuse and glossy lobes), the camera's importance (the pixel-area
Jacobian of SampleResponse), the geometric / shading cosine corrections of
light-subpath BSDF samples (a sphere with interpolated normals) and the
(1 / eta)^2 radiance scaling of transmission on the light subpath (a rough
glass sphere). A dropped term or a swapped cosine shifts the light tracer's
image by far more than the bars."""
import numpy as np
import pytest

import oracle_lib as ol
from seesharp_amd import abi, scenes
from seesharp_amd.scene import (GenericMaterial, GlossyEmitter, Mesh)

BATCHES = 8


def spheres_box(width=32, height=32, glossy_light=False):
    """The box with a rough glass sphere and a smooth-normal metal sphere."""
    s = scenes.box_scene(width, height)
    v, f = scenes.icosphere(2)
    for centre, radius, params in (
            ((0.32, 0.7, 0.6), 0.2, dict(BaseColor=(0.9, 0.9, 0.9), Roughness=0.25, Metallic=0.0,
                                           SpecularTransmittance=0.9, IndexOfRefraction=1.5)),
            ((0.72, 0.55, 0.35), 0.12, dict(BaseColor=(0.8, 0.6, 0.3), Roughness=0.4, Metallic=1.0,
                                             SpecularTransmittance=0.0, IndexOfRefraction=1.5))):
        m = Mesh((v * radius + np.array(centre)).astype(np.float32), f.reshape(-1), v.astype(np.float32))
        m.Material = GenericMaterial(GenericMaterial.Parameters(SpecularTintStrength=1.0, Anisotropic=0.0,
                                                                **params))
        s.Meshes.append(m)
    if glossy_light:
        light = s.Emitters[0].Mesh
        s.Emitters = GlossyEmitter.MakeFromMesh(light, (15.0, 15.0, 15.0), 8.0)
    return s


def tilted_floor_box(width=32, height=32):
    """The box whose floor's shading normals are tilted ~25 degrees off its face normal."""
    s = scenes.box_scene(width, height)
    floor = s.Meshes[0]
    n = np.array([0.35, 1.0, 0.25], np.float32)
    floor.ShadingNormals = np.tile(n / np.linalg.norm(n), (floor.Vertices.shape[0], 1)).astype(np.float32)
    return s


def batch_images(o, kind, cfg, iters):
    out = []
    for b in range(BATCHES):
        if kind == "lt":
            cfg.base_seed_light = 0x13C0FEFE + 7919 * b
            img, _, nans = o.render_light_tracer(cfg, 0, iters)
        else:
            cfg.base_seed = 0xC030114 + 7919 * b
            img, _, nans = o.render(cfg, 0, iters)
        assert nans == 0
        out.append(img.astype(np.float64))
    return np.stack(out)


def z_scores(a, b, block=8):
    """z-scores of the block means of two batch stacks (B, H, W, 3)."""
    B, H, W, _ = a.shape

    def blocks(x):
        return x.reshape(B, H // block, block, W // block, block, 3).mean(axis=(2, 4, 5))

    ba, bb = blocks(a), blocks(b)
    se = np.sqrt(ba.var(axis=0, ddof=1) / B + bb.var(axis=0, ddof=1) / B)
    return (ba.mean(axis=0) - bb.mean(axis=0)) / np.maximum(se, 1e-12)


def compare(scene, max_depth, min_depth=2, iters=48):
    o = ol.OracleScene(scene)
    lt = batch_images(o, "lt", ol.lt_settings(max_depth=max_depth, min_depth=min_depth), iters)
    pt = batch_images(o, "pt", ol.settings(max_depth=max_depth, min_depth=min_depth), iters)
    m_lt, m_pt = lt.mean(axis=(1, 2, 3)), pt.mean(axis=(1, 2, 3))
    z_mean = (m_lt.mean() - m_pt.mean()) / np.sqrt(m_lt.var(ddof=1) / BATCHES + m_pt.var(ddof=1) / BATCHES)
    z = z_scores(lt, pt)
    print(f"mean LT {m_lt.mean():.5f} PT {m_pt.mean():.5f} z {z_mean:.2f}; block |z| max {np.abs(z).max():.2f}")
    return m_lt.mean(), m_pt.mean(), z_mean, z


@pytest.mark.parametrize("max_depth", [2, 5])
def test_light_tracer_matches_path_tracer_on_box(max_depth):
    lt, pt, z_mean, z = compare(scenes.box_scene(32, 32), max_depth)
    assert pt > 0
    assert abs(z_mean) < 4.0
    assert np.abs(z).max() < 5.0


def test_light_tracer_matches_path_tracer_glass_and_metal():
    # (a light subpath without the (1 / eta)^2 transmission scaling is off by ~10 sigma here)
    lt, pt, z_mean, z = compare(spheres_box(), 6, iters=160)
    assert abs(z_mean) < 4.0
    assert np.abs(z).max() < 5.0


def test_light_tracer_matches_path_tracer_tilted_shading_normals():
    # (without the geometric / shading cosine corrections of the light subpath the
    # floor's light-traced image is off by many sigma)
    lt, pt, z_mean, z = compare(tilted_floor_box(), 4, iters=96)
    assert abs(z_mean) < 4.0
    assert np.abs(z).max() < 5.0


def test_light_tracer_matches_path_tracer_glossy_emitter():
    lt, pt, z_mean, z = compare(spheres_box(glossy_light=True), 4)
    assert abs(z_mean) < 4.0
    assert np.abs(z).max() < 5.0


def test_light_tracer_min_depth():
    lt, pt, z_mean, z = compare(scenes.box_scene(32, 32), 5, min_depth=3)
    assert abs(z_mean) < 4.0
    assert np.abs(z).max() < 5.0


def test_light_tracer_depth_one_is_black():
    # MaxDepth 1: a light path has no vertex, so nothing is splatted
    o = ol.OracleScene(scenes.box_scene(16, 16))
    img, st, nans = o.render_light_tracer(ol.lt_settings(max_depth=1), 0, 2)
    assert not img.any() and st.num_rays == 0


def test_light_tracer_running_mean_split():
    # 2 + 2 iterations continue the running mean of a 4-iteration render (same arithmetic)
    o = ol.OracleScene(scenes.box_scene(16, 16))
    cfg = ol.lt_settings()
    full, _, _ = o.render_light_tracer(cfg, 0, 4)
    part, _, _ = o.render_light_tracer(cfg, 0, 2)
    part, _, _ = o.render_light_tracer(cfg, 2, 2, image=part)
    assert np.array_equal(full, part)


def test_light_tracer_with_background_is_vertex_cache_bidir():
    # with a background the configuration also traces camera paths that hit it
    # (BidirBase.OnBackgroundHit): VertexCacheBidir without connections, next
    # events or emitter hits, camera paths from the default BaseSeedCamera
    from test_oracle_bidir_background import open_box

    o = ol.OracleScene(open_box(24, 24))
    lt, st_lt, _ = o.render_light_tracer(ol.lt_settings(max_depth=5), 0, 2)
    vcb, st_v, _ = o.render_vcm(ol.vcm_settings(algorithm=abi.VERTEX_CACHE_BIDIR, base_seed_camera=0xC030114,
                                                max_depth=5, num_connections=0, num_shadow_rays=0,
                                                enable_hitting=False), 0, 2)
    assert lt.mean() > 0 and st_lt.num_rays == st_v.num_rays
    np.testing.assert_array_equal(lt, vcb)
