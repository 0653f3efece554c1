"""GPU tests of the plain single-kernel tracer (SSPT_EXEC_SINGLE_KERNEL): it
evaluates the same per-vertex code and RNG draws as the wavefront, one thread
per path, so both render bit-identical images with identical ray counts."""
import numpy as np
import pytest

import oracle_lib as ol
from seesharp_amd import GpuBatch, GpuScene, abi, scenes

pytestmark = pytest.mark.gpu


def render_both(device, g, cfg, spp):
    out = []
    for single in (False, True):
        device.set_execution_mode(single)
        try:
            g.clear()
            st = g.render(cfg, 0, spp)
            out.append((g.read_image(), st))
        finally:
            device.set_execution_mode(False)
    return out


@pytest.mark.parametrize("which", ["box", "large"])
def test_single_kernel_renders_the_wavefront_image(device, which):
    if which == "box":
        s, cfg, spp = scenes.box_scene(64, 64), ol.settings(total_spp=8, max_depth=5), 8
    else:
        s = scenes.large_scene(96, 64, num_spheres=40, grid=48, ground_textures=32)
        cfg, spp = ol.settings(total_spp=6, max_depth=8), 6
    g = GpuScene(device, s)
    (wf, swf), (sk, ssk) = render_both(device, g, cfg, spp)
    assert wf.mean() > 0
    assert np.array_equal(wf, sk)
    for k in ("num_rays", "num_shadow_rays", "num_shadow_rays_traced", "num_occluded", "num_ray_hits"):
        assert swf[k] == ssk[k], k


def test_single_kernel_batch_matches_wavefront_batch(device):
    sc = [scenes.box_scene(48, 40), scenes.large_scene(64, 48, num_spheres=20, grid=32)]
    cfgs = [ol.settings(total_spp=4, max_depth=5), ol.settings(total_spp=3, max_depth=8)]
    b = GpuBatch(device, sc)
    b.render(cfgs)
    wf = [b.read_image(i) for i in range(2)]
    device.set_execution_mode(True)
    try:
        b.render(cfgs)
        sk = [b.read_image(i) for i in range(2)]
    finally:
        device.set_execution_mode(False)
    for a, r in zip(sk, wf):  # per-scene sums: fp32 atomics in any order
        assert float(np.max(np.abs(a - r))) <= 1e-5 * max(1.0, float(np.abs(r).max()))


def test_adjoint_requires_the_wavefront(device):
    s = scenes.box_scene(32, 32, textured=True, tex_res=4)
    g = GpuScene(device, s)
    device.set_execution_mode(True)
    try:
        with pytest.raises(abi.SsptError, match="wavefront"):
            g.render_adjoint(ol.settings(total_spp=1, max_depth=3), np.ones((32, 32, 3), np.float32), 0, 1)
    finally:
        device.set_execution_mode(False)
