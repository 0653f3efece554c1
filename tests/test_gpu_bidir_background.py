"""GPU ClassicBidir, VertexCacheBidir and VCM on scenes with a background
(light paths from EnvironmentMap.SampleRay, background hits under
EmitterHitMis, background next events) against the oracle at the same seeds.

Every light and camera path draws the oracle's random numbers; the images
agree up to the order of the fp32 splat additions and the few ulps of the GPU's
arithmetic (hardware reciprocal, sine and cosine: the background ray's disc
point moves by ulps, which the weights pass on smoothly). Scenes: the open box
under a sky map with and without its ceiling emitter (without one, every light
path starts at the background), the open box with flat-shaded glass and metal
spheres, and the reduced large scene with its sky and smooth-normal spheres.
Settings cover MaxDepth 1 and 2 (a background path stores its first hit
whatever MaxDepth), NumShadowRays 2, MinDepth 3 and each technique switched
off. Observed on MI355X: relMSE 3e-14 .. 1.7e-7 for the scenes (the large
scene 1.3e-11), up to 8.7e-7 for the settings, where one pixel in a thousand
takes a path that a last-ulp difference diverts; identical ray counts. Bars:
relMSE at about 10x the observed value, at most 1 % of the pixels off by more
than rtol 1e-3, image means within 1e-3 and ray counts within 1e-3."""
import numpy as np
import pytest

import oracle_lib as ol
from seesharp_amd import GpuScene, abi, scenes
from test_oracle_bidir_background import open_box

pytestmark = pytest.mark.gpu
VCB = abi.VERTEX_CACHE_BIDIR


def relmse(x, r):
    return float(np.mean((x - r) ** 2 / (r ** 2 + 1e-2)))


def compare(g, o, kind, cfg, iters, bar, name):
    g.clear()
    sg = getattr(g, kind)(cfg, 0, iters)
    img_g = g.read_image()
    img_o, so, nan_o = getattr(o, kind)(cfg, 0, iters)
    err = relmse(img_g, img_o)
    div = float(np.mean(np.any(~np.isclose(img_g, img_o, rtol=1e-3, atol=1e-5), axis=-1)))
    print(f"{kind} {name}: mean gpu {img_g.mean():.6f} oracle {img_o.mean():.6f} relMSE {err:.2e} "
          f"diverted {div:.2e} rays {sg['num_rays']} / {so.num_rays} shadow {sg['num_shadow_rays']} / "
          f"{so.num_shadow_rays}", flush=True)
    assert img_o.mean() > 0
    assert g.nan_count() == nan_o == 0
    assert err <= bar
    assert div <= 0.01
    assert abs(img_g.mean() - img_o.mean()) <= 1e-3 * img_o.mean()
    assert abs(sg["num_rays"] - so.num_rays) <= 1e-3 * so.num_rays
    assert abs(sg["num_shadow_rays"] - so.num_shadow_rays) <= 1e-3 * so.num_shadow_rays


CASES = {  # scene, MaxDepth, relMSE bar
    "sky_and_emitter": (lambda: open_box(64, 64), 5, 2e-6),
    "sky_only": (lambda: open_box(64, 48, emitter=False), 5, 1e-6),
    "spheres": (lambda: open_box(48, 64, spheres=True), 6, 1e-6),
    "large_with_sky": (lambda: scenes.large_scene(96, 96, num_spheres=60, grid=64), 6, 1e-6),
}
RENDERS = {"bidir": ("render_bidir", ol.bidir_settings, {}),
           "vertex_cache_bidir": ("render_vcm", ol.vcm_settings, dict(algorithm=VCB)),
           "vcm": ("render_vcm", ol.vcm_settings, dict(algorithm=abi.VCM))}


@pytest.mark.parametrize("render", list(RENDERS))
@pytest.mark.parametrize("name", list(CASES))
def test_background_matches_oracle(device, name, render):
    make, depth, bar = CASES[name]
    kind, settings, kw = RENDERS[render]
    s = make()
    compare(GpuScene(device, s), ol.OracleScene(s), kind, settings(max_depth=depth, **kw), 2, bar, name)


BIDIR_SETTINGS = [dict(max_depth=1), dict(max_depth=2), dict(max_depth=6, num_shadow_rays=2),
                  dict(max_depth=8, min_depth=3), dict(max_depth=6, enable_connections=False),
                  dict(max_depth=6, enable_light_tracer=False)]
VCM_SETTINGS = [dict(max_depth=1), dict(max_depth=2), dict(max_depth=6, num_shadow_rays=2),
                dict(max_depth=8, min_depth=3), dict(max_depth=6, num_connections=3),
                dict(max_depth=6, enable_hitting=False), dict(max_depth=6, num_light_paths=300),
                dict(max_depth=6, merge_primary=True, max_num_photons=1),
                dict(max_depth=6, disable_correl_aware_mis=True)]


def ident(kw):
    return "-".join(f"{k}={v}" for k, v in kw.items())


@pytest.mark.parametrize("kw", BIDIR_SETTINGS, ids=ident)
def test_background_bidir_settings_match_oracle(device, kw):
    s = open_box(48, 48, spheres=True)
    compare(GpuScene(device, s), ol.OracleScene(s), "render_bidir", ol.bidir_settings(**kw), 2, 1e-5, ident(kw))


@pytest.mark.parametrize("algorithm", [VCB, abi.VCM], ids=["vertex_cache_bidir", "vcm"])
@pytest.mark.parametrize("kw", VCM_SETTINGS, ids=ident)
def test_background_vcm_settings_match_oracle(device, kw, algorithm):
    s = open_box(48, 48, spheres=True)
    compare(GpuScene(device, s), ol.OracleScene(s), "render_vcm", ol.vcm_settings(algorithm=algorithm, **kw), 2,
            1e-5, ident(kw))


def textured_glossy_open_box(width=48, height=48):
    """The open box under the sky with textured back wall and floor (albedo and
    roughness textures), a glossy ceiling emitter and the flat-shaded spheres."""
    from seesharp_amd.scene import GlossyEmitter

    s = open_box(width, height, spheres=True)
    back, floor = scenes.textured_materials(32, 5, None)
    s.Meshes[0].Material = floor
    s.Meshes[1].Material = back  # the ceiling was removed: index 1 is the back wall
    light = s.Emitters[0].Mesh
    s.Emitters = GlossyEmitter.MakeFromMesh(light, (15.0, 15.0, 15.0), 6.0)
    return s


@pytest.mark.parametrize("render", list(RENDERS))
def test_background_with_textures_and_glossy_emitter_matches_oracle(device, render):
    kind, settings, kw = RENDERS[render]
    s = textured_glossy_open_box()
    compare(GpuScene(device, s), ol.OracleScene(s), kind, settings(max_depth=6, **kw), 2, 1e-5, "textured glossy")
