"""GPU tests of the path-replay adjoint (sspt_render_adjoint) and the torch
autograd wrapper.

Albedo: with Metallic 0, SpecularTintStrength 0 and SpecularTransmittance 0
the GenericMaterial's lobe selection and pdfs do not depend on the base
colour, a DiffuseMaterial's never do, and Russian roulette is off up to depth
4. For a fixed seed the rendered image is then a polynomial in the albedo
texels, so a same-seed central difference is the exact derivative of the
estimate the adjoint differentiates: the two agree to the FD truncation and
fp32 rounding (2% bound, observed far below).

Roughness moves the GGX sample directions, which the adjoint holds fixed, so
the same-seed difference and the adjoint agree only in expectation: that test
compares both against their Monte Carlo standard errors.
"""
import numpy as np
import pytest
import torch

import oracle_lib as ol
from seesharp_amd import GpuScene, abi, autograd, scenes

pytestmark = pytest.mark.gpu


def textured_box(diffuse: bool, res: int = 96, tex_res: int = 8):
    s = scenes.box_scene(res, res, textured=True, tex_res=tex_res, spec_tint=0.0, diffuse_textures=diffuse)
    return s


def loss_of(img, w):
    return float(np.sum(img.astype(np.float64) * w))


def render_loss(g, cfg, spp, w, first=0):
    """Loss of the mean image of iterations first..first+spp-1."""
    return loss_of(g.render_mean(cfg, first, spp), w)


def test_albedo_gradients_match_central_differences_on_64_texels(device):
    """The differentiation workload: the box at 512^2 with 256^2 albedo and
    roughness textures on the back wall and floor (initial value 0.5), L2 loss
    against the render of smooth-noise textures, forward and adjoint 32 spp on
    the same paths; central differences with h = 1e-3 on 64 random albedo texels
    (56 among the texels the loss depends on: within 2%; 8 among the rest: within
    0.1% of the largest derivative checked). MaxDepth 4 and SpecularTintStrength 0 keep the
    same-seed estimate smooth in the albedo (deeper paths: Russian roulette
    decisions, and with tint the lobe choice, depend on the albedo and would make
    it jump); the oracle-tangent test covers MaxDepth 6 and tinted materials."""
    res, tex_res, spp = 512, 256, 32
    target_scene = scenes.box_scene(res, res, textured=True, tex_res=tex_res, spec_tint=0.0)
    tg = GpuScene(device, target_scene)
    cfg = ol.settings(total_spp=spp, max_depth=4)
    tg.clear()
    tg.render(cfg, 0, 256)
    target = tg.read_image().astype(np.float64)
    tg.close()
    s = scenes.box_scene(res, res, textured=True, tex_res=tex_res, spec_tint=0.0, uniform_textures=0.5)
    g = GpuScene(device, s)
    npx = res * res * 3

    def loss():
        img = g.render_mean(cfg, 1000, spp).astype(np.float64)
        return float(np.sum((img - target) ** 2) / npx), img

    _, img = loss()
    g.clear_gradients()
    g.render_adjoint(cfg, (2.0 * (img - target) / npx).astype(np.float32), 1000, spp)
    rng = np.random.default_rng(2024)
    albedo = [g.texture_index(m.MaterialParameters.BaseColor) for m in (s.Meshes[0].Material, s.Meshes[2].Material)]
    grads = {k: g.gradient(k).astype(np.float64) for k in albedo}
    # 28 texels per texture among those the loss depends on (gradient >= 2% of
    # the texture's largest), 4 per texture among the others
    picks = []
    for idx in albedo:
        a = np.abs(grads[idx])
        strong = np.argwhere(a >= 0.02 * a.max())
        weak = np.argwhere(a < 0.02 * a.max())
        assert len(strong) >= 28 and len(weak) >= 4
        picks += [(idx, tuple(strong[i]), True) for i in rng.choice(len(strong), 28, replace=False)]
        picks += [(idx, tuple(weak[i]), False) for i in rng.choice(len(weak), 4, replace=False)]
    h = 1e-3
    rows = []
    for idx, (y, x, c), strong in picks:
        base = s.DeviceTextures[idx].Image
        vals = []
        for sign in (1.0, -1.0):
            t = base.copy()
            t[y, x, c] += sign * h
            g.write_texture(idx, t)
            vals.append(loss()[0])
        g.write_texture(idx, base)
        rows.append(((vals[0] - vals[1]) / (2 * h), grads[idx][y, x, c], strong))
    assert len(rows) == 64
    big = max(abs(r[0]) for r in rows)
    assert big > 0
    bad = [r for r in rows if r[2] and abs(r[1] - r[0]) > 0.02 * abs(r[0])]
    bad += [r for r in rows if not r[2] and abs(r[1] - r[0]) > 1e-3 * big]
    assert not bad, bad


@pytest.mark.parametrize("kind", ["generic_albedo", "diffuse_albedo", "roughness"])
def test_adjoint_matches_oracle_tangent_on_the_same_paths(device, kind):
    """GPU gradient against the oracle's tangent (same seeds, every sampling
    decision held fixed, BSDF values by central differences of the oracle's own
    material evaluation) at MaxDepth 6 with tinted GenericMaterials: random
    directions and single texels. The GPU and the oracle follow the same paths
    except where a last-ulp difference changes a sampling decision or a
    traversal tie, hence 1%."""
    s = scenes.box_scene(48, 48, textured=True, tex_res=8, diffuse_textures=(kind == "diffuse_albedo"))
    g = GpuScene(device, s)
    o = ol.OracleScene(s)
    spp = 8
    cfg = ol.settings(total_spp=spp, max_depth=6)
    if kind == "roughness":
        tex = s.Meshes[0].Material.MaterialParameters.Roughness
    else:
        tex = s.Meshes[2].Material.MaterialParameters.BaseColor
    idx = g.texture_index(tex)
    rng = np.random.default_rng(17)
    w = rng.uniform(0.5, 1.5, (48, 48, 3)).astype(np.float32)
    g.clear_gradients()
    g.render_adjoint(cfg, w, 0, spp)
    grad = g.gradient(idx).astype(np.float64)
    assert np.all(np.isfinite(grad))
    directions = [rng.normal(size=grad.shape).astype(np.float32) for _ in range(2)]
    for y, x in ((2, 3), (5, 5), (6, 1)):
        e = np.zeros(grad.shape, np.float32)
        e[y, x] = 1.0
        directions.append(e)
    seen = 0
    for v in directions:
        tan, nans = o.render_tangent(cfg, 0, spp, idx, v, eps=1e-3)
        assert nans == 0
        ref = float(np.sum(w.astype(np.float64) * tan))
        scale = float(np.sum(np.abs(w.astype(np.float64) * tan)))
        got = float(np.sum(grad * v))
        if scale == 0.0:  # a texel no path reaches: no gradient either
            assert abs(got) <= 1e-6 * float(np.abs(grad).max()), (kind, got)
            continue
        seen += 1
        assert abs(got - ref) <= 0.01 * scale, (kind, got, ref, scale)
    assert seen >= 4


def test_gradient_is_linear_in_dloss_dimage(device):
    s = textured_box(False, res=64)
    g = GpuScene(device, s)
    cfg = ol.settings(total_spp=4, max_depth=5)
    rng = np.random.default_rng(3)
    w1 = rng.uniform(-1, 1, (64, 64, 3)).astype(np.float32)
    w2 = rng.uniform(-1, 1, (64, 64, 3)).astype(np.float32)
    n = len(s.DeviceTextures)

    def grads(w):
        g.clear_gradients()
        g.render_adjoint(cfg, w, 0, 4)
        return [g.gradient(k).astype(np.float64) for k in range(n)]

    a, b = grads(w1), grads(w2)
    c = grads((2.0 * w1 - 0.5 * w2).astype(np.float32))
    for k in range(n):
        ref = 2.0 * a[k] - 0.5 * b[k]
        scale = max(1e-6, float(np.abs(ref).max()))
        assert float(np.abs(c[k] - ref).max()) <= 1e-4 * scale, k


def test_zero_dloss_gives_zero_gradient_and_clear_resets(device):
    s = textured_box(False, res=48)
    g = GpuScene(device, s)
    cfg = ol.settings(total_spp=2, max_depth=5)
    g.clear_gradients()
    g.render_adjoint(cfg, np.ones((48, 48, 3), np.float32), 0, 2)
    assert any(np.abs(g.gradient(k)).max() > 0 for k in range(len(s.DeviceTextures)))
    g.clear_gradients()
    g.render_adjoint(cfg, np.zeros((48, 48, 3), np.float32), 0, 2)
    for k in range(len(s.DeviceTextures)):
        assert np.all(g.gradient(k) == 0.0)


def test_untextured_scene_rejects_gradient_reads(device):
    g = GpuScene(device, scenes.box_scene(32, 32))
    cfg = ol.settings(total_spp=1, max_depth=3)
    g.clear_gradients()
    g.render_adjoint(cfg, np.ones((32, 32, 3), np.float32), 0, 1)
    buf = np.zeros(16, np.float32)
    st = device.lib.sspt_gradient_read(device.ptr, g.ptr, 0, abi.fptr(buf))
    assert st == abi.INVALID_ARGUMENT
    assert "texture index out of range" in device.lib.sspt_last_error(device.ptr).decode()


def test_roughness_gradient_matches_central_difference_in_expectation(device):
    """d/dh of the loss when every roughness texel of the floor grows by h."""
    s = textured_box(False, res=64, tex_res=8)
    g = GpuScene(device, s)
    spp, blocks = 64, 8
    cfg = ol.settings(total_spp=spp, max_depth=4)
    w = np.ones((64, 64, 3))
    idx = g.texture_index(s.Meshes[0].Material.MaterialParameters.Roughness)
    base = s.DeviceTextures[idx].Image.copy()
    h = 0.05
    adj, fd = [], []
    for b in range(blocks):
        first = b * spp
        g.clear_gradients()
        g.render_adjoint(cfg, w.astype(np.float32), first + 100000, spp)
        adj.append(float(g.gradient(idx).astype(np.float64).sum()))
        vals = []
        for sign in (1.0, -1.0):
            g.write_texture(idx, base + sign * h)
            vals.append(render_loss(g, cfg, spp, w, first))
        g.write_texture(idx, base)
        fd.append((vals[0] - vals[1]) / (2 * h))
    adj, fd = np.array(adj), np.array(fd)
    se = np.sqrt(adj.var(ddof=1) / blocks + fd.var(ddof=1) / blocks)
    diff = abs(adj.mean() - fd.mean())
    assert diff <= 4.0 * se + 0.02 * abs(fd.mean()), (adj.mean(), fd.mean(), se)


def test_autograd_function_matches_native_adjoint(device):
    s = textured_box(False, res=48)
    g = GpuScene(device, s)
    cfg = ol.settings(total_spp=4, max_depth=5)
    idx = g.texture_index(s.Meshes[2].Material.MaterialParameters.BaseColor)
    tex = torch.tensor(s.DeviceTextures[idx].Image, device="cuda:0", requires_grad=True)
    w = torch.rand(48, 48, 3, generator=torch.Generator().manual_seed(1)).to("cuda:0")
    img = autograd.render(g, cfg, [(idx, tex)], first_iteration=0, backward_first_iteration=50)
    assert img.shape == (48, 48, 3) and img.device.type == "cuda"
    (img * w).sum().backward()
    g.clear_gradients()
    g.render_adjoint(cfg, w.cpu().numpy(), 50, 4)
    ref = g.gradient(idx)
    got = tex.grad.cpu().numpy()
    assert np.abs(got - ref).max() <= 1e-4 * max(1e-6, float(np.abs(ref).max()))


def test_albedo_optimisation_recovers_target_albedo(device):
    """Gradient steps on the back-wall albedo (start 0.2) with noisy 16-spp L2
    losses recover the albedo (0.7) that rendered the target: the texels move to
    the target and a low-noise loss drops. (The 16-spp loss itself is dominated
    by its Monte Carlo noise floor near the optimum.)"""
    s = textured_box(True, res=48, tex_res=4)
    g = GpuScene(device, s)
    cfg = ol.settings(total_spp=16, max_depth=4)
    idx = g.texture_index(s.Meshes[2].Material.MaterialParameters.BaseColor)
    target_tex = np.full_like(s.DeviceTextures[idx].Image, 0.7)
    g.write_texture(idx, target_tex)
    target = g.render_mean(cfg, 100000, 256).astype(np.float64)

    def clean_loss(t):
        g.write_texture(idx, t)
        return float(np.mean((g.render_mean(cfg, 200000, 256).astype(np.float64) - target) ** 2))

    tex = torch.full(target_tex.shape, 0.2, requires_grad=True)
    loss0 = clean_loss(tex.detach().numpy())
    opt = torch.optim.Adam([tex], lr=0.08)
    tgt = torch.tensor(target, dtype=torch.float32)
    for it in range(16):
        opt.zero_grad()
        img = autograd.render(g, cfg, [(idx, tex)], first_iteration=it * 32, backward_first_iteration=it * 32 + 16)
        ((img - tgt) ** 2).mean().backward()
        opt.step()
        with torch.no_grad():
            tex.clamp_(0.0, 1.0)
    final = tex.detach().numpy()
    loss1 = clean_loss(final)
    assert float(np.median(np.abs(final - 0.7))) < 0.15, final
    assert loss1 < 0.5 * loss0, (loss0, loss1)


def test_device_dloss_matches_host_dloss(device):
    """dLoss/dImage passed by its device pointer (a torch tensor on the GPU, copied
    device to device) gives the gradients of the same values passed from host memory
    (up to the order of the fp32 gradient atomics)."""
    s = textured_box(diffuse=False, res=64, tex_res=16)
    g = GpuScene(device, s)
    cfg = ol.settings(total_spp=4, max_depth=5)
    rng = np.random.default_rng(3)
    dimg = rng.normal(0.0, 1.0, (64, 64, 3)).astype(np.float32)
    grads = []
    for src in (dimg, torch.from_numpy(dimg).to("cuda:0")):
        g.clear_gradients()
        g.render_adjoint(cfg, src, 0, 4)
        grads.append([g.gradient(i) for i in range(len(g.texture_shapes()))])
    for a, b in zip(*grads):
        assert np.abs(a).max() > 0
        np.testing.assert_allclose(b, a, rtol=1e-4, atol=1e-6 * np.abs(a).max())
