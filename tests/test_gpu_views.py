"""GPU tests of view sets (sspt_views_create): many cameras on one scene in one
world, rendered and differentiated in shared waves, with every view's
gradients added into the one set of texture gradients."""
import numpy as np
import pytest

import oracle_lib as ol
from seesharp_amd import GpuBatch, GpuScene, scenes

pytestmark = pytest.mark.gpu


def relmse(x, r):
    return float(np.mean((x - r) ** 2 / (r ** 2 + 1e-2)))


def small_textured_scene():
    return scenes.large_scene(64, 48, num_spheres=20, grid=32, ground_textures=16)


def test_views_match_single_camera_renders(device):
    """View i of a view set draws the same RNG streams as a render of the scene
    with camera i: images agree up to fp32 summation order."""
    s = small_textured_scene()
    cams = scenes.circle_views(4, radius=8.0, height=2.0)
    views = GpuBatch.views(device, s, cams)
    cfgs = [ol.settings(total_spp=n, max_depth=8) for n in (4, 2, 4, 3)]
    views.render(cfgs)
    for i, cam in enumerate(cams):
        s.Camera = cam
        g = GpuScene(device, s)
        g.clear()
        g.render(cfgs[i], 0, cfgs[i].total_spp)
        single = g.read_image()
        assert single.mean() > 0
        assert relmse(views.read_image(i), single) < 1e-8, i


def test_views_adjoint_is_the_sum_of_single_view_adjoints(device):
    s = small_textured_scene()
    cams = scenes.circle_views(3, radius=8.0, height=2.0)
    views = GpuBatch.views(device, s, cams)
    cfgs = [ol.settings(total_spp=3, max_depth=6)] * 3
    rng = np.random.default_rng(9)
    dl = [rng.uniform(-1.0, 1.0, (48, 64, 3)).astype(np.float32) for _ in cams]
    views.clear_gradients()
    views.render_adjoint(cfgs, dl, first_iteration=7)
    idx = [views.texture_index(t) for t in s.DeviceTextures]
    got = [views.gradient(k).astype(np.float64) for k in idx]
    ref = [np.zeros_like(x) for x in got]
    for i, cam in enumerate(cams):
        s.Camera = cam
        g = GpuScene(device, s)
        g.clear_gradients()
        g.render_adjoint(cfgs[i], dl[i], 7, 3)
        for k in range(len(idx)):
            ref[k] += g.gradient(k)
    for a, b in zip(got, ref):
        scale = float(np.abs(b).max())
        assert scale > 0
        assert float(np.abs(a - b).max()) <= 1e-4 * scale


def test_views_gradient_matches_oracle_tangent(device):
    """The summed gradient of all views against the oracle's same-path tangent
    of each view (1%: the GPU and the oracle follow the same paths except where
    a last-ulp difference changes a sampling decision)."""
    s = small_textured_scene()
    cams = scenes.circle_views(2, radius=8.0, height=2.0)
    views = GpuBatch.views(device, s, cams)
    spp = 4
    cfgs = [ol.settings(total_spp=spp, max_depth=8)] * 2
    rng = np.random.default_rng(21)
    w = [rng.uniform(0.5, 1.5, (48, 64, 3)).astype(np.float32) for _ in cams]
    views.clear_gradients()
    views.render_adjoint(cfgs, w, first_iteration=0)
    for tex in s.DeviceTextures:
        idx = views.texture_index(tex)
        grad = views.gradient(idx).astype(np.float64)
        v = rng.normal(size=grad.shape).astype(np.float32)
        ref = 0.0
        scale = 0.0
        for i, cam in enumerate(cams):
            s.Camera = cam
            o = ol.OracleScene(s)
            tan, nans = o.render_tangent(cfgs[i], 0, spp, idx, v, eps=1e-3)
            assert nans == 0
            ref += float(np.sum(w[i].astype(np.float64) * tan))
            scale += float(np.sum(np.abs(w[i].astype(np.float64) * tan)))
        got = float(np.sum(grad * v))
        assert scale > 0
        assert abs(got - ref) <= 0.01 * scale, (idx, got, ref, scale)
