"""GPU tests of the denoiser feature layers (albedo / normal at the primary hit)."""
import numpy as np
import pytest

import oracle_lib as ol
from seesharp_amd import GpuBatch, GpuPathTracer, GpuScene, scenes

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("which", ["box", "large"])
def test_layers_match_oracle(device, which):
    if which == "box":
        s = scenes.box_scene(48, 48, textured=True, tex_res=8)
    else:
        s = scenes.large_scene(64, 48, num_spheres=30, grid=32, ground_textures=16)
    g = GpuScene(device, s)
    cfg = ol.settings(total_spp=4, max_depth=5, enable_denoiser=True)
    g.clear()
    g.render(cfg, 0, 4)
    o = ol.OracleScene(s)
    alb_o, nrm_o = o.primary_aov(cfg, 0, 4)
    alb, nrm = g.read_layer("albedo"), g.read_layer("normal")
    assert alb.max() > 0 and np.abs(nrm).max() > 0
    assert np.abs(alb - alb_o).max() <= 1e-6
    assert np.abs(nrm - nrm_o).max() <= 1e-6


def test_layers_split_iterations_single_kernel_and_off(device):
    s = scenes.large_scene(64, 48, num_spheres=30, grid=32, ground_textures=16)
    g = GpuScene(device, s)
    cfg = ol.settings(total_spp=4, max_depth=6, enable_denoiser=True)
    g.clear()
    g.render(cfg, 0, 4)
    ref = (g.read_layer("albedo"), g.read_layer("normal"))
    g.clear()
    g.render(cfg, 0, 2)
    g.render(cfg, 2, 2)
    assert np.array_equal(g.read_layer("albedo"), ref[0]) and np.array_equal(g.read_layer("normal"), ref[1])
    device.set_execution_mode(True)
    try:
        g.clear()
        g.render(cfg, 0, 4)
        assert np.array_equal(g.read_layer("albedo"), ref[0]) and np.array_equal(g.read_layer("normal"), ref[1])
    finally:
        device.set_execution_mode(False)
    g.clear()
    g.render(ol.settings(total_spp=4, max_depth=6), 0, 4)
    assert np.all(g.read_layer("albedo") == 0) and np.all(g.read_layer("normal") == 0)


def test_batch_layers_match_single_scene_layers(device):
    sc = [scenes.box_scene(40, 32, textured=True, tex_res=4), scenes.large_scene(48, 40, num_spheres=20, grid=24)]
    cfgs = [ol.settings(total_spp=3, max_depth=5, enable_denoiser=True),
            ol.settings(total_spp=2, max_depth=8, enable_denoiser=True)]
    b = GpuBatch(device, sc)
    b.render(cfgs)
    for i, s in enumerate(sc):
        g = GpuScene(device, s)
        g.clear()
        g.render(cfgs[i], 0, cfgs[i].total_spp)
        for name in ("albedo", "normal"):
            a, r = b.read_layer(i, name), g.read_layer(name)
            assert np.abs(a - r).max() <= 1e-6 * max(1.0, float(np.abs(r).max())), (i, name)


def test_path_tracer_writes_framebuffer_layers(device):
    s = scenes.box_scene(32, 32)
    pt = GpuPathTracer(device)
    pt.TotalSpp = 2
    pt.MaxDepth = 3
    pt.Render(s)
    assert set(s.FrameBuffer.Layers) == {"albedo", "normal"}
    assert s.FrameBuffer.Layers["albedo"].shape == (32, 32, 3)
