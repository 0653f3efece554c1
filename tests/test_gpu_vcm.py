"""GPU VertexCacheBidir and VertexConnectionAndMerging (sspt_render_vcm) against
the oracle at the same seeds.

Every light and camera path draws the oracle's random numbers, every
connection picks the oracle's cache vertex and every merge the oracle's
photons, so the images agree up to the order of the fp32 splat additions and
the few ulps of the GPU's shading arithmetic. A last-ulp difference can move
a photon across the merge radius or divert a path, so a few pixels may differ
visibly: bars as for the light tracer and ClassicBidir, relMSE and the share
of pixels off by more than rtol 1e-3. The observed values are printed."""
import numpy as np
import pytest

import oracle_lib as ol
from seesharp_amd import GpuBatch, GpuScene, VertexCacheBidir, VertexConnectionAndMerging, abi, scenes
from seesharp_amd.scene import EnvironmentMap
from test_oracle_bidir import glossy_floor_box
from test_oracle_light_tracer import spheres_box, tilted_floor_box

pytestmark = pytest.mark.gpu
VCB = abi.VERTEX_CACHE_BIDIR


def relmse(x, r):
    return float(np.mean((x - r) ** 2 / (r ** 2 + 1e-2)))


def compare(g, o, cfg, iters, bar, name, div_bar=0.01):
    g.clear()
    sg = g.render_vcm(cfg, 0, iters)
    img_g = g.read_image()
    img_o, so, nan_o = o.render_vcm(cfg, 0, iters)
    err = relmse(img_g, img_o)
    div = float(np.mean(np.any(~np.isclose(img_g, img_o, rtol=1e-3, atol=1e-5), axis=-1)))
    print(f"{name}: mean gpu {img_g.mean():.6f} oracle {img_o.mean():.6f} relMSE {err:.2e} diverted {div:.2e} "
          f"rays {sg['num_rays']} / {so.num_rays} shadow {sg['num_shadow_rays']} / {so.num_shadow_rays}", flush=True)
    assert img_o.mean() > 0
    assert g.nan_count() == nan_o == 0
    assert err <= bar
    assert div <= div_bar
    assert abs(img_g.mean() - img_o.mean()) <= 1e-3 * img_o.mean()
    assert abs(sg["num_rays"] - so.num_rays) <= 1e-3 * so.num_rays
    assert abs(sg["num_shadow_rays"] - so.num_shadow_rays) <= 1e-3 * so.num_shadow_rays


SCENES = {  # scene, MaxDepth, relMSE bar
    "box": (lambda: scenes.box_scene(64, 64), 5, 1e-7),
    "glass_metal": (lambda: spheres_box(64, 64), 6, 1e-7),
    "glossy_light": (lambda: spheres_box(64, 48, glossy_light=True), 4, 1e-7),
    "tilted_normals": (lambda: tilted_floor_box(48, 64), 5, 1e-7),
    "glossy_floor": (lambda: glossy_floor_box(48, 48), 5, 1e-7),
    "large_no_env": (lambda: scenes.large_scene(96, 96, num_spheres=60, grid=64, with_env=False), 6, 1e-5),
}


@pytest.mark.parametrize("algorithm", [VCB, abi.VCM], ids=["vertex_cache_bidir", "vcm"])
@pytest.mark.parametrize("name", list(SCENES))
def test_vcm_matches_oracle(device, name, algorithm):
    make, depth, bar = SCENES[name]
    s = make()
    compare(GpuScene(device, s), ol.OracleScene(s), ol.vcm_settings(algorithm=algorithm, max_depth=depth), 3, bar,
            f"{name} {algorithm}")


SETTINGS = [dict(num_connections=0), dict(num_connections=3), dict(num_light_paths=500),
            dict(enable_hitting=False), dict(enable_light_tracer=False), dict(num_shadow_rays=0),
            dict(num_shadow_rays=2), dict(max_depth=8, min_depth=3), dict(merge_primary=True),
            dict(max_num_photons=1), dict(max_num_photons=32), dict(disable_correl_aware_mis=True),
            dict(enable_merging=False), dict(max_depth=1), dict(max_depth=12),
            dict(base_seed_camera=12345, base_seed_light=999)]


@pytest.mark.parametrize("kw", SETTINGS, ids=lambda kw: "-".join(f"{k}={v}" for k, v in kw.items()))
def test_vcm_settings_match_oracle(device, kw):
    s = spheres_box(48, 48)
    kw = dict(kw)
    kw.setdefault("max_depth", 6)
    compare(GpuScene(device, s), ol.OracleScene(s), ol.vcm_settings(**kw), 2, 1e-7, str(kw))


def grazing_floor(width=64, height=64):
    """A 40 x 40 diffuse floor seen from 5 cm above it towards the horizon, a glossy
    box on it, lit by an area light: the primary hits near the horizon have pixel
    footprints of a few up to millions of photon-grid cells, so the photon query
    takes its shell search, its early stop and its scan of every photon."""
    from seesharp_amd.scene import (DiffuseEmitter, DiffuseMaterial, FrameBuffer, GenericMaterial, Mesh,
                                    PerspectiveCamera, Scene, create_look_at)

    s = Scene()

    def add(v, idx, mat):
        m = Mesh(v, idx)
        m.Material = mat
        s.Meshes.append(m)
        return m

    add(*scenes.quad((-20, 0, -20), (-20, 0, 20), (20, 0, 20), (20, 0, -20)),
        DiffuseMaterial(DiffuseMaterial.Parameters(BaseColor=(0.7, 0.7, 0.7))))
    add(*scenes.box_mesh((0.0, 0.25, -3.0), (0.5, 0.5, 0.5), 30.0),
        GenericMaterial(GenericMaterial.Parameters(BaseColor=(0.8, 0.5, 0.3), Roughness=0.3, Metallic=0.5,
                                                   SpecularTintStrength=1.0, Anisotropic=0.0,
                                                   SpecularTransmittance=0.0, IndexOfRefraction=1.5)))
    light = add(*scenes.light_quad(0.0, 3.0, -2.0, 1.5), DiffuseMaterial(DiffuseMaterial.Parameters(
        BaseColor=(0.0, 0.0, 0.0))))
    s.Emitters.extend(DiffuseEmitter.MakeFromMesh(light, (20.0, 20.0, 20.0)))
    s.Camera = PerspectiveCamera(create_look_at((0.0, 0.05, 4.0), (0.0, 0.05, -10.0), (0, 1, 0)), 50.0)
    s.FrameBuffer = FrameBuffer(width, height)
    return s


@pytest.mark.parametrize("kw", [dict(), dict(merge_primary=True), dict(merge_primary=True, max_num_photons=32)],
                         ids=["default", "merge_primary", "merge_primary-32_photons"])
def test_vcm_grazing_footprints_match_oracle(device, kw):
    s = grazing_floor()
    compare(GpuScene(device, s), ol.OracleScene(s), ol.vcm_settings(max_depth=5, **kw), 2, 1e-7, str(kw))


def test_vertex_cache_bidir_reduces_to_light_tracer(device):
    s = spheres_box(48, 48)
    g = GpuScene(device, s)
    g.clear()
    g.render_vcm(ol.vcm_settings(algorithm=VCB, max_depth=6, num_connections=0, num_shadow_rays=0,
                                 enable_hitting=False), 0, 3)
    a = g.read_image()
    g.clear()
    g.render_light_tracer(ol.lt_settings(max_depth=6), 0, 3)
    b = g.read_image()
    assert b.mean() > 0
    np.testing.assert_allclose(a, b, rtol=1e-5, atol=1e-7)


def test_vcm_running_mean_continues(device):
    s = scenes.box_scene(32, 32)
    g = GpuScene(device, s)
    cfg = ol.vcm_settings(max_depth=5)
    g.clear()
    g.render_vcm(cfg, 0, 4)
    full = g.read_image()
    g.clear()
    g.render_vcm(cfg, 0, 2)
    g.render_vcm(cfg, 2, 2)
    np.testing.assert_allclose(g.read_image(), full, rtol=1e-5, atol=1e-7)


def test_vcm_integrators(device):
    s = scenes.box_scene(32, 32)
    o = ol.OracleScene(s)
    for cls, alg in ((VertexCacheBidir, VCB), (VertexConnectionAndMerging, abi.VCM)):
        it = cls(device)
        it.MaxDepth = 5
        it.NumIterations = 2
        fb = it.Render(s)
        ref, _, _ = o.render_vcm(ol.vcm_settings(algorithm=alg, max_depth=5), 0, 2)
        assert fb.CurIteration == 2
        assert relmse(fb.Image, ref) <= 1e-7


def test_vcm_unsupported(device):
    s = scenes.box_scene(16, 16)
    g = GpuScene(device, s)
    for bad in (dict(max_num_photons=0), dict(max_num_photons=33), dict(algorithm=5), dict(num_connections=-1)):
        with pytest.raises(abi.SsptError):
            g.render_vcm(ol.vcm_settings(**bad), 0, 1)
    s.Background = EnvironmentMap(np.ones((8, 16, 3), np.float32))  # a background is drawn from
    GpuScene(device, s).render_vcm(ol.vcm_settings(), 0, 1)
    batch = GpuBatch(device, [scenes.box_scene(16, 16), scenes.box_scene(16, 16)])
    with pytest.raises(abi.SsptError):
        device.check(device.lib.sspt_render_vcm(device.ptr, batch.ptr, ol.vcm_settings(), 0, 1, None))
