"""The oracle's bidirectional integrators on scenes with a background
(BidirBase.Light.cs:103-133,174-189,283-299; BidirBase.Camera.cs:338-418,532-561;
RandomWalk.cs:89-128; EnvironmentMap.cs:87-130; Scene.cs:114-141).

With a background, SelectLight picks it with probability 1 / (1 + |emitters|)
and traces the light path from a ray of EnvironmentMap.SampleRay (a map
direction and a point on the scene-spanning disc), next-event estimation picks
it with the same probability, and camera paths that leave the scene add the
background's radiance under EmitterHitMis. The techniques' weights still sum to
one for every path, so ClassicBidir and VertexCacheBidir stay unbiased
estimators of the path tracer's image: compared through z-scores of image and
8x8 block means over independent seed batches, on an open box lit by a sky map
and a ceiling emitter, the same box lit by the sky alone (every light path
starts at the background), and the box with flat-shaded glass and metal
spheres. VCM differs from itself without merging by the photon-mapping bias.

Three behaviours of the reference restated as they are, and kept out of the
unbiasedness checks: the first hit of a background light path continues with
the BSDF sample's weight without the light subpath's shading-normal
corrections (RandomWalk.cs:117-127), so the scenes use face normals; the
background's next-event pdf carries no NumShadowRays factor in
OnBackgroundHit and PerformNextEventEstimation while NextEventPdf does
(ClassicBidir.cs:42-44), so NumShadowRays stays 1; and background hits skip
the MinDepth and EnableHitting tests while VertexCacheBidir's NextEventMis
leaves the hits out without EnableHitting, so MinDepth stays 1 and hitting on.

Planted in the oracle on the CPU, each of these mistakes fails a check here: the
background ancestor's reverse pdf converted to area measure like a surface's,
the background's selection probability dropped from the light-path pdf, and
the first background vertex's pdf converted by the area Jacobian."""
import numpy as np
import pytest

import oracle_lib as ol
from seesharp_amd import abi, scenes
from seesharp_amd.scene import EnvironmentMap, GenericMaterial, Mesh
from test_oracle_bidir import check

BATCHES = 8


def soft_sky(width=32, height=16):
    """A small lat-long sky: horizon-to-zenith gradient, dim ground, a broad sun."""
    rows = (np.arange(height) + 0.5) / height * np.pi
    cols = (np.arange(width) + 0.5) / width * 2 * np.pi
    theta, phi = np.meshgrid(rows, cols, indexing="ij")
    d = np.stack([np.sin(theta) * np.cos(phi), np.cos(theta), np.sin(theta) * np.sin(phi)], -1)
    up = np.clip(d[..., 1], 0, 1)[..., None]
    sky = np.array([0.9, 0.85, 0.8]) * (1 - up) + np.array([0.3, 0.5, 1.0]) * up
    img = np.where(d[..., 1:2] >= 0, sky, np.array([0.2, 0.17, 0.12]))
    sun = np.array([0.5, 0.7, 0.5])
    sun /= np.linalg.norm(sun)
    img = img + np.clip((d @ sun - 0.9) / 0.1, 0, 1)[..., None] * np.array([6.0, 5.5, 5.0])
    return img.astype(np.float32)


def open_box(width=32, height=32, emitter=True, spheres=False):
    """The box without its ceiling under the sky map; the ceiling emitter stays
    unless `emitter` is false; `spheres` adds a flat-shaded rough glass and a
    metal sphere."""
    s = scenes.box_scene(width, height)
    del s.Meshes[1]  # the ceiling
    if not emitter:
        light = s.Emitters[0].Mesh
        s.Emitters = []
        s.Meshes = [m for m in s.Meshes if m is not light]
    if spheres:
        v, f = scenes.icosphere(2)
        for centre, radius, params in (
                ((0.32, 0.7, 0.6), 0.2, dict(BaseColor=(0.9, 0.9, 0.9), Roughness=0.25, Metallic=0.0,
                                               SpecularTransmittance=0.9, IndexOfRefraction=1.5)),
                ((0.72, 0.55, 0.35), 0.12, dict(BaseColor=(0.8, 0.6, 0.3), Roughness=0.4, Metallic=1.0,
                                                 SpecularTransmittance=0.0, IndexOfRefraction=1.5))):
            m = Mesh((v * radius + np.array(centre)).astype(np.float32), f.reshape(-1))
            m.Material = GenericMaterial(GenericMaterial.Parameters(SpecularTintStrength=1.0, Anisotropic=0.0,
                                                                    **params))
            s.Meshes.append(m)
    s.Background = EnvironmentMap(soft_sky())
    return s


SCENES = {  # scene, MaxDepth, iterations per batch
    "sky_and_emitter": (lambda: open_box(), 5, 32),
    "sky_only": (lambda: open_box(emitter=False), 5, 32),
    "spheres": (lambda: open_box(spheres=True), 6, 48),
}


def pt_batches(o, iters, max_depth):
    out = []
    for b in range(BATCHES):
        img, _, nans = o.render(ol.settings(base_seed=0xC030114 + 7919 * b, max_depth=max_depth), 0, iters)
        assert nans == 0
        out.append(img.astype(np.float64))
    return np.stack(out)


def batches(render, settings, iters, seed=0, **kw):
    out = []
    for b in range(BATCHES):
        cfg = settings(base_seed_camera=0x5EED0000 + 7919 * b + seed, base_seed_light=0x13C0FEFE + 7919 * b + seed,
                       **kw)
        img, _, nans = render(cfg, 0, iters)
        assert nans == 0
        out.append(img.astype(np.float64))
    return np.stack(out)


@pytest.mark.parametrize("name", list(SCENES))
def test_bidir_with_background_matches_path_tracer(name):
    make, depth, iters = SCENES[name]
    o = ol.OracleScene(make())
    check(batches(o.render_bidir, ol.bidir_settings, iters, max_depth=depth), pt_batches(o, iters, depth))


@pytest.mark.parametrize("name", list(SCENES))
def test_vertex_cache_bidir_with_background_matches_path_tracer(name):
    make, depth, iters = SCENES[name]
    o = ol.OracleScene(make())
    vcb = batches(o.render_vcm, ol.vcm_settings, iters, algorithm=abi.VERTEX_CACHE_BIDIR, max_depth=depth)
    check(vcb, pt_batches(o, iters, depth))


@pytest.mark.parametrize("variant", [dict(enable_connections=False), dict(enable_light_tracer=False),
                                     dict(enable_connections=False, enable_light_tracer=False)],
                         ids=["no_connections", "no_light_tracer", "hits_and_next_events"])
def test_bidir_with_background_techniques_agree(variant):
    o = ol.OracleScene(open_box(spheres=True))
    full = batches(o.render_bidir, ol.bidir_settings, 48, max_depth=6)
    part = batches(o.render_bidir, ol.bidir_settings, 48, seed=0x51DE, max_depth=6, **variant)
    check(full, part)


@pytest.mark.parametrize("variant", [dict(num_connections=0), dict(num_connections=3),
                                     dict(enable_light_tracer=False), dict(num_light_paths=300),
                                     dict(num_connections=0, enable_light_tracer=False)],
                         ids=lambda v: "-".join(f"{k}={x}" for k, x in v.items()))
def test_vertex_cache_bidir_with_background_techniques_agree(variant):
    o = ol.OracleScene(open_box(spheres=True))
    kw = dict(algorithm=abi.VERTEX_CACHE_BIDIR, max_depth=6)
    check(batches(o.render_vcm, ol.vcm_settings, 48, **kw),
          batches(o.render_vcm, ol.vcm_settings, 48, seed=0x51DE, **kw, **variant))


@pytest.mark.parametrize("name", ["sky_and_emitter", "sky_only"])
def test_vcm_with_background_merging_bias(name):
    make, depth, iters = SCENES[name]
    o = ol.OracleScene(make())
    diffs, means = [], []
    for b in range(BATCHES):
        kw = dict(base_seed_camera=0x5EED0000 + 7919 * b, base_seed_light=0x13C0FEFE + 7919 * b, max_depth=depth)
        full, _, _ = o.render_vcm(ol.vcm_settings(**kw), 0, 16)
        plain, _, _ = o.render_vcm(ol.vcm_settings(enable_merging=False, **kw), 0, 16)
        diffs.append(float((full.astype(np.float64) - plain).mean()))
        means.append(float(plain.mean()))
    rel = float(np.mean(diffs) / np.mean(means))
    print(f"{name}: merging changes the image mean by {rel:+.4f}")
    assert -0.012 < rel < 0.002


def test_background_light_paths_are_drawn():
    # with the sky alone every light path starts at the background: the light tracer
    # alone (no connections, hits or next events would see anything else) still
    # lights the box interior
    o = ol.OracleScene(open_box(emitter=False))
    img, st, _ = o.render_vcm(ol.vcm_settings(algorithm=abi.VERTEX_CACHE_BIDIR, num_connections=0,
                                              num_shadow_rays=0, enable_hitting=False), 0, 4)
    assert img.mean() > 0.05 and st.num_rays > 0
