"""GPU ClassicBidir (sspt_render_bidir: ClassicBidir.cs over BidirBase.*.cs,
BidirPathPdfs.cs, RandomWalk.cs) against the oracle at the same seeds, and
against the GPU path tracer as an independent estimator of the same image.

Same seeds: every light and camera path draws the oracle's random numbers, so
the images agree up to the order of the fp32 splat additions (atomics on the
GPU) and the few ulps of the GPU's shading arithmetic (hardware reciprocal and
square root), which the balance-heuristic weights pass on smoothly. Bars at
about 10x or more of the observed relMSE, as for the light tracer: <= 1e-8 for
the box scenes, <= 1e-5 for the reduced large scene with its smooth-normal
spheres, with at most 1 % of the pixels off by more than rtol 1e-3. The ray
counts agree except for the rare path that a last-ulp difference diverts."""
import numpy as np
import pytest

import oracle_lib as ol
from seesharp_amd import ClassicBidir, GpuBatch, GpuScene, abi, scenes
from seesharp_amd.scene import EnvironmentMap
from test_oracle_bidir import glossy_floor_box
from test_oracle_light_tracer import spheres_box, tilted_floor_box

pytestmark = pytest.mark.gpu


def relmse(x, r):
    return float(np.mean((x - r) ** 2 / (r ** 2 + 1e-2)))


CASES = {  # scene, MaxDepth, relMSE bar
    "box": (lambda: scenes.box_scene(64, 64), 5, 1e-8),
    "glass_metal": (lambda: spheres_box(64, 64), 6, 1e-8),
    "glossy_light": (lambda: spheres_box(64, 48, glossy_light=True), 4, 1e-8),
    "tilted_normals": (lambda: tilted_floor_box(48, 64), 5, 1e-8),
    "glossy_floor": (lambda: glossy_floor_box(48, 48), 5, 1e-8),
    "large_no_env": (lambda: scenes.large_scene(96, 96, num_spheres=60, grid=64, with_env=False), 6, 1e-5),
}


def compare(g, o, cfg, iters, bar, name):
    g.clear()
    sg = g.render_bidir(cfg, 0, iters)
    img_g = g.read_image()
    img_o, so, nan_o = o.render_bidir(cfg, 0, iters)
    err = relmse(img_g, img_o)
    div = float(np.mean(np.any(~np.isclose(img_g, img_o, rtol=1e-3, atol=1e-5), axis=-1)))
    print(f"{name}: mean gpu {img_g.mean():.6f} oracle {img_o.mean():.6f} relMSE {err:.2e} diverted {div:.2e} "
          f"rays {sg['num_rays']} / {so.num_rays} shadow {sg['num_shadow_rays']} / {so.num_shadow_rays}", flush=True)
    assert img_o.mean() > 0
    assert g.nan_count() == nan_o == 0
    assert err <= bar
    assert div <= 0.01
    assert abs(img_g.mean() - img_o.mean()) <= 1e-3 * img_o.mean()
    assert abs(sg["num_rays"] - so.num_rays) <= 1e-3 * so.num_rays
    assert abs(sg["num_shadow_rays"] - so.num_shadow_rays) <= 1e-3 * so.num_shadow_rays
    assert abs(sg["num_ray_hits"] - so.num_ray_hits) <= 1e-3 * so.num_ray_hits


@pytest.mark.parametrize("name", list(CASES))
def test_bidir_matches_oracle(device, name):
    make, depth, bar = CASES[name]
    s = make()
    compare(GpuScene(device, s), ol.OracleScene(s), ol.bidir_settings(max_depth=depth), 3, bar, name)


SETTINGS = [dict(max_depth=8, min_depth=3), dict(max_depth=4, num_shadow_rays=0),
            dict(max_depth=5, num_shadow_rays=3), dict(max_depth=6, enable_connections=False),
            dict(max_depth=6, enable_light_tracer=False), dict(max_depth=2, base_seed_camera=12345,
                                                                 base_seed_light=999),
            dict(max_depth=1), dict(max_depth=12, num_light_paths=48 * 48)]


@pytest.mark.parametrize("kw", SETTINGS, ids=lambda kw: "-".join(f"{k}={v}" for k, v in kw.items()))
def test_bidir_settings_match_oracle(device, kw):
    s = spheres_box(48, 48)
    compare(GpuScene(device, s), ol.OracleScene(s), ol.bidir_settings(**kw), 2, 1e-8, str(kw))


def test_bidir_running_mean_continues(device):
    s = scenes.box_scene(32, 32)
    g = GpuScene(device, s)
    cfg = ol.bidir_settings(max_depth=5)
    g.clear()
    g.render_bidir(cfg, 0, 4)
    full = g.read_image()
    g.clear()
    g.render_bidir(cfg, 0, 2)
    g.render_bidir(cfg, 2, 2)
    split = g.read_image()
    np.testing.assert_allclose(split, full, rtol=1e-5, atol=1e-7)


def test_bidir_agrees_with_path_tracer(device):
    # independent estimators of the same image
    s = spheres_box(64, 64)
    g = GpuScene(device, s)
    batches = 8
    bd, pt = [], []
    for b in range(batches):
        g.clear()
        g.render_bidir(ol.bidir_settings(base_seed_camera=0x5EED0000 + 7919 * b,
                                         base_seed_light=0x13C0FEFE + 7919 * b, max_depth=6), 0, 128)
        bd.append(g.read_image().astype(np.float64))
        g.clear()
        g.render(ol.settings(base_seed=0xC030114 + 7919 * b, max_depth=6), 0, 128)
        pt.append(g.read_image().astype(np.float64))
    bd, pt = np.stack(bd), np.stack(pt)

    def blocks(x):
        return x.reshape(batches, 8, 8, 8, 8, 3).mean(axis=(2, 4, 5))

    bb, bp = blocks(bd), blocks(pt)
    se = np.sqrt(bb.var(axis=0, ddof=1) / batches + bp.var(axis=0, ddof=1) / batches)
    z = (bb.mean(axis=0) - bp.mean(axis=0)) / np.maximum(se, 1e-12)
    mb, mp = bd.mean(axis=(1, 2, 3)), pt.mean(axis=(1, 2, 3))
    z_mean = (mb.mean() - mp.mean()) / np.sqrt(mb.var(ddof=1) / batches + mp.var(ddof=1) / batches)
    print(f"BD {mb.mean():.5f} PT {mp.mean():.5f} z {z_mean:.2f} block |z| max {np.abs(z).max():.2f}")
    assert abs(z_mean) < 4.0
    assert np.abs(z).max() < 5.0


def test_bidir_integrator(device):
    s = scenes.box_scene(32, 32)
    bd = ClassicBidir(device)
    bd.MaxDepth = 5
    bd.NumIterations = 3
    fb = bd.Render(s)
    img_o, _, _ = ol.OracleScene(s).render_bidir(ol.bidir_settings(max_depth=5), 0, 3)
    assert fb.CurIteration == 3
    assert relmse(fb.Image, img_o) <= 1e-8


def test_bidir_unsupported(device):
    s = scenes.box_scene(16, 16)
    g = GpuScene(device, s)
    with pytest.raises(abi.SsptError):  # ClassicBidir.cs:72-75
        g.render_bidir(ol.bidir_settings(num_light_paths=100), 0, 1)
    s.Background = EnvironmentMap(np.ones((8, 16, 3), np.float32))  # a background is drawn from
    GpuScene(device, s).render_bidir(ol.bidir_settings(), 0, 1)
    batch = GpuBatch(device, [scenes.box_scene(16, 16), scenes.box_scene(16, 16)])
    with pytest.raises(abi.SsptError):
        device.check(device.lib.sspt_render_bidir(device.ptr, batch.ptr, ol.bidir_settings(), 0, 1, None))
