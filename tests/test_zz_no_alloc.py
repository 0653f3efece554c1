"""sspt_render_noalloc (through the C ABI) against sspt_render.

The allocation-free entry point renders with every wave buffer in the library's
per-GPU scratch array, so a wave holds only ~172 000 path slots: a 1024^2 image
goes in waves of part of an iteration, a small image in waves of whole
iterations. Both must give sspt_render's results: the same running-mean image,
denoiser layers and outlier cache bit for bit, and the same ray counters
(one shadow ray per kind, so every per-path sum is added in a fixed order).
Two calls in a row, the second on a smaller scene, check that the scratch
contents a call leaves behind do not leak into the next.
"""
import ctypes as C

import numpy as np
import pytest

import oracle_lib as ol
from seesharp_amd import GpuScene, abi, scenes

pytestmark = pytest.mark.gpu

COUNTERS = ("num_rays", "num_shadow_rays", "num_occluded", "num_ray_hits", "num_shadow_rays_traced")


def rendered(g, cfg, first, num, noalloc):
    g.clear()
    st = (g.render_noalloc if noalloc else g.render)(cfg, first, num)
    return g.read_image(), st


def test_noalloc_matches_render_large_then_small(device):
    large = scenes.large_scene(1024, 1024)
    gl = GpuScene(device, large)
    assert large.FrameBuffer.Width * large.FrameBuffer.Height > 200_000  # waves of part of an iteration
    cfg = ol.settings(total_spp=4, max_depth=8)
    print("noalloc: large scene 1024^2, 4 iterations", flush=True)
    ref, st_ref = rendered(gl, cfg, 0, 4, False)
    img, st = rendered(gl, cfg, 0, 4, True)
    assert np.array_equal(img, ref)
    assert all(st[k] == st_ref[k] for k in COUNTERS), (st, st_ref)
    assert st["num_iterations"] == 4
    gl.close()
    # the second call, on a smaller scene: waves of whole iterations
    box = scenes.box_scene(64, 48)
    gb = GpuScene(device, box)
    cfg = ol.settings(total_spp=8, max_depth=5)
    print("noalloc: box 64x48, 8 iterations", flush=True)
    ref, st_ref = rendered(gb, cfg, 0, 8, False)
    img, st = rendered(gb, cfg, 0, 8, True)
    assert np.array_equal(img, ref)
    assert all(st[k] == st_ref[k] for k in COUNTERS), (st, st_ref)


def test_noalloc_split_calls_continue_the_running_mean(device):
    s = scenes.large_scene(640, 480, num_spheres=100, grid=128)
    g = GpuScene(device, s)
    cfg = ol.settings(total_spp=6, max_depth=8)
    print("noalloc: 640x480 split 2 + 4 iterations", flush=True)
    ref, _ = rendered(g, cfg, 0, 6, False)
    g.clear()
    g.render_noalloc(cfg, 0, 2)
    g.render_noalloc(cfg, 2, 4)
    assert np.array_equal(g.read_image(), ref)


def test_noalloc_layers_and_outliers_match_render(device):
    s = scenes.large_scene(512, 512, num_spheres=100, grid=128)
    g = GpuScene(device, s)
    g.track_outliers(3)
    cfg = ol.settings(total_spp=5, max_depth=8, enable_denoiser=True)
    print("noalloc: 512x512 with denoiser layers and 3 outliers", flush=True)
    ref, _ = rendered(g, cfg, 0, 5, False)
    ref_alb, ref_nrm, ref_out = g.read_layer("albedo"), g.read_layer("normal"), g.read_outliers()
    img, _ = rendered(g, cfg, 0, 5, True)
    assert np.array_equal(img, ref)
    assert np.array_equal(g.read_layer("albedo"), ref_alb)
    assert np.array_equal(g.read_layer("normal"), ref_nrm)
    out = g.read_outliers()
    for a, b in zip(out, ref_out):
        assert np.array_equal(a, b)


def test_noalloc_rejects_what_it_cannot_do(device):
    s = scenes.box_scene(32, 32)
    g = GpuScene(device, s)
    lib = device.lib
    stats = abi.RayStats()
    budget = ol.settings(total_spp=4, max_render_time_ms=10)
    assert lib.sspt_render_noalloc(device.ptr, g.ptr, C.byref(budget), 0, 4, C.byref(stats)) == 5  # UNSUPPORTED
    too_many = ol.settings(total_spp=4, num_shadow_rays=100000)
    assert lib.sspt_render_noalloc(device.ptr, g.ptr, C.byref(too_many), 0, 4, C.byref(stats)) == 1  # INVALID
    assert b"scratch" in lib.sspt_last_error(device.ptr)
