"""GPU tests of the batched path (many scenes in one world) against the single-scene path."""
import numpy as np
import pytest

import oracle_lib as ol
from seesharp_amd import GpuBatch, GpuScene, scenes

pytestmark = pytest.mark.gpu


def relmse(x, r):
    return float(np.mean((x - r) ** 2 / (r ** 2 + 1e-2)))


def test_batch_matches_single_scene_renders(device):
    """Each scene of a batch draws the same RNG streams as its own render: the
    images agree up to fp32 summation order."""
    sc = [scenes.box_scene(64, 64), scenes.box_scene(96, 80), scenes.large_scene(64, 48, num_spheres=20, grid=32)]
    spp = [8, 4, 16]
    cfgs = [ol.settings(total_spp=n, max_depth=d) for n, d in zip(spp, (5, 3, 8))]
    b = GpuBatch(device, sc)
    b.render(cfgs)
    for i, s in enumerate(sc):
        g = GpuScene(device, s)
        g.clear()
        g.render(cfgs[i], 0, spp[i])
        single = g.read_image()
        batched = b.read_image(i)
        assert relmse(batched, single) < 1e-8, i
        assert abs(float(batched.mean()) - float(single.mean())) <= 1e-5 * float(single.mean())


def test_batch_with_small_waves_matches(device):
    """Waves that split iterations of a scene give the same images."""
    sc = [scenes.box_scene(40, 40), scenes.box_scene(24, 32)]
    cfgs = [ol.settings(total_spp=6, max_depth=5), ol.settings(total_spp=3, max_depth=4)]
    b = GpuBatch(device, sc)
    device.set_wave_capacity(1000)
    try:
        b.render(cfgs)
        small = [b.read_image(i) for i in range(2)]
    finally:
        device.set_wave_capacity(0)
    b.render(cfgs)
    for i in range(2):
        assert relmse(small[i], b.read_image(i)) < 1e-8


def test_graph_and_direct_launches_agree(device):
    s = scenes.box_scene(64, 64)
    g = GpuScene(device, s)
    cfg = ol.settings(total_spp=8, max_depth=5)
    device.set_launch_config(0, 0, False)
    try:
        g.clear()
        g.render(cfg, 0, 8)
        direct = g.read_image()
    finally:
        device.set_launch_config(0, 0, True)
    for _ in range(3):  # replays of the captured wave, with eager work in between
        g.clear()
        g.render(cfg, 0, 8)
        assert np.array_equal(g.read_image(), direct)


def test_persistent_and_full_grids_render_identical_images(device):
    """Grid sizes only change which block walks which queue chunk: every path
    computes the same values either way (one shadow ray of each kind per vertex,
    so the per-slot accumulation order is fixed too)."""
    s = scenes.large_scene(96, 64, num_spheres=30, grid=48)
    g = GpuScene(device, s)
    cfg = ol.settings(total_spp=8, max_depth=8)
    g.clear()
    g.render(cfg, 0, 8)
    full = g.read_image()
    resident = device.launch_config()
    for tb, sb in ((resident["trace_per_cu"] * 256, resident["shade_per_cu"] * 256), (37, 5)):
        device.set_launch_config(tb, sb, True)
        try:
            g.clear()
            g.render(cfg, 0, 8)
            assert np.array_equal(g.read_image(), full), (tb, sb)
        finally:
            device.set_launch_config(0, 0, True)



def test_block_shapes_render_identical_images(device):
    """Every prebuilt block shape (traversal 64 / 128 / 256 threads, forward shade
    128 / 256 / 512) renders the default shape's image bit for bit: a block only
    shares LDS stack space and queue appends, never a path's arithmetic."""
    s = scenes.large_scene(96, 64, num_spheres=30, grid=48)
    g = GpuScene(device, s)
    cfg = ol.settings(total_spp=6, max_depth=8)
    g.clear()
    g.render(cfg, 0, 6)
    ref = g.read_image()
    try:
        for tb, sb in ((64, 128), (256, 512), (128, 512), (64, 256)):
            device.set_block_shape(tb, sb)
            assert device.block_shape() == (tb, sb)
            g.clear()
            g.render(cfg, 0, 6)
            assert np.array_equal(g.read_image(), ref), (tb, sb)
    finally:
        device.set_block_shape(0, 0)
    assert device.block_shape() == (128, 256)


def test_autotuner_pick_renders_identical_images(device):
    """The autotuner (block shape, wave capacity, grids by coordinate descent) leaves
    its pick set on the device; a render with it equals the default render bit for bit."""
    from seesharp_amd import autotune

    s = scenes.large_scene(128, 96, num_spheres=40, grid=64)
    g = GpuScene(device, s)
    cfg = ol.settings(total_spp=8, max_depth=8)
    g.clear()
    g.render(cfg, 0, 8)
    ref = g.read_image()

    def run():
        g.clear()
        g.render(cfg, 0, 8)

    try:
        best, results = autotune.tune(device, run, capacities=(1 << 16, 1 << 18, 1 << 20))
        assert len(results) == 9 + 3 + 9
        assert device.block_shape() == (best["trace_block"], best["shade_block"])
        g.clear()
        g.render(cfg, 0, 8)
        assert np.array_equal(g.read_image(), ref), best
    finally:
        device.set_block_shape(0, 0)
        device.set_wave_capacity(0)
        device.set_launch_config(0, 0, True)
