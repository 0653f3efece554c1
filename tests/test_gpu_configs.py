"""The batching and differentiation configurations at their full sizes, on the GPU.

* C2, batching: the 64 scenes of seeds 1000..1063 (1.8k-194k triangles, 128^2-512^2,
  4/16/64 spp, MaxDepth 8) in one world. Every scene's image equals its own
  single-scene render to fp32 summation order and the ray counters of the batch
  are the sums of the single renders'; four scenes against the oracle at the same
  seeds on seeded pixel subsets.
* C4, differentiation + batching: the large scene with 1024^2 RGB albedo and
  roughness textures on the ground, 32 views on a circle at 512^2, 16 spp,
  MaxDepth 8. Three views' forward images against the oracle on pixel subsets;
  the batched adjoint of all 32 views (dLoss/dImage non-zero on pixels of two
  views whose paths the GPU and the oracle follow alike) against the oracle's
  same-path tangent summed over those views.
* C3, differentiation at MaxDepth 6: the 512^2 textured box with 256^2 textures,
  32 + 32 spp. (a) Central differences (h = 1e-3) of the L2 loss linearised at the
  expected image on 64 albedo texels against the adjoint, paired per seed over 24
  decorrelated seeds: beyond depth 4 Russian roulette depends on the albedo, so a
  same-seed difference carries the 1/q terms and jumps where a decision flips;
  the seed average of FD - adjoint is compared with its standard error. (b) The adjoint against the 4096-spp CPU reference gradient
  (the oracle's tangent, whose expectation the adjoint estimates without bias)
  along random directions of the albedo and roughness textures, the GPU averaged
  over 32 decorrelated 32-spp passes, within 4 combined standard errors.

Same-seed image bars as the large-scene parity test (glossy GenericMaterials turn
ulp differences into diverted paths): relMSE <= 3e-4 and at most 3 % of the
pixels off by more than rtol 1e-3.
"""
import numpy as np
import pytest

import oracle_lib as ol
from seesharp_amd import GpuBatch, GpuScene, abi, scenes

pytestmark = pytest.mark.gpu

SEED = 0xC030114


def relmse(x, r):
    return float(np.mean((x - r) ** 2 / (r ** 2 + 1e-2)))


def diverted(x, r):
    return float(np.mean(np.any(~np.isclose(x, r, rtol=1e-3, atol=1e-5), axis=-1)))


def subset(rng, w, h, n):
    return np.sort(rng.choice(w * h, min(n, w * h), replace=False)).astype(np.int32)


# ---------------------------------------------------------------------- C2

def test_c2_batch_of_64_scenes(device):
    sc, spp = scenes.batch_scenes(64, 1000)
    cfgs = [ol.settings(total_spp=n, max_depth=8) for n in spp]
    b = GpuBatch(device, sc)
    print(f"C2: {b.info()}", flush=True)
    st = b.render(cfgs)
    sums = {k: 0 for k in ("num_rays", "num_shadow_rays", "num_ray_hits", "num_shadow_rays_traced",
                           "num_occluded")}
    worst = 0.0
    for i, s in enumerate(sc):
        g = GpuScene(device, s)
        g.clear()
        one = g.render(cfgs[i], 0, spp[i])
        for k in sums:
            sums[k] += one[k]
        err = relmse(b.read_image(i), g.read_image())
        worst = max(worst, err)
        assert err < 1e-8, (i, err)
        g.close()
    print(f"C2: worst batch-vs-single relMSE {worst:.2e}; counters {st}")
    for k in sums:
        assert st[k] == sums[k], k
    # four scenes against the oracle: fewest / most triangles, most samples, one more
    tris = [s.NumTriangles for s in sc]
    picks = sorted({int(np.argmin(tris)), int(np.argmax(tris)), int(np.argmax(spp)), 37})
    rng = np.random.default_rng(2)
    for i in picks:
        s = sc[i]
        w, h = s.FrameBuffer.Width, s.FrameBuffer.Height
        px = subset(rng, w, h, 2048)
        img_o, _, nans = ol.OracleScene(s).render(cfgs[i], 0, spp[i], pixels=px)
        ref = img_o.reshape(-1, 3)[px]
        got = b.read_image(i).reshape(-1, 3)[px]
        err, div = relmse(got, ref), diverted(got, ref)
        print(f"C2 scene {i} ({tris[i]} tris, {w}x{h}, {spp[i]} spp): relMSE {err:.2e}, diverted {div:.2e}")
        assert nans == 0
        assert err <= 3e-4 and div <= 0.03, i


# ---------------------------------------------------------------------- C4

@pytest.fixture(scope="module")
def c4(device):
    res = 512
    s = scenes.large_scene(res, res, ground_textures=1024, ground_uniform=None, ground_texture_seed=5)
    cams = scenes.circle_views(32, radius=8.0, height=2.0)
    views = GpuBatch.views(device, s, cams)
    return s, cams, views


def test_c4_forward_views_match_oracle(c4):
    s, cams, views = c4
    cfgs = [ol.settings(total_spp=16, max_depth=8)] * len(cams)
    views.render(cfgs)
    rng = np.random.default_rng(4)
    for v in (0, 13, 27):
        s.Camera = cams[v]
        px = subset(rng, 512, 512, 2048)
        img_o, _, nans = ol.OracleScene(s).render(cfgs[v], 0, 16, pixels=px)
        ref = img_o.reshape(-1, 3)[px]
        got = views.read_image(v).reshape(-1, 3)[px]
        err, div = relmse(got, ref), diverted(got, ref)
        print(f"C4 view {v}: relMSE {err:.2e}, diverted {div:.2e}")
        assert nans == 0
        assert err <= 3e-4 and div <= 0.03, v


def test_c4_batched_adjoint_matches_oracle_tangent(c4):
    """dLoss/dImage is non-zero on seeded pixels of views 5 and 21 and zero on every
    other pixel and view; the adjoint of all 32 views (one batched launch)
    contracted with a random direction of each ground texture equals the sum over
    the two views of the oracle's same-path tangent on those pixels, within 1 % of
    the contraction's absolute sum. The pixels are those of 2048 seeded ones whose
    forward estimate agrees with the oracle's to rtol 1e-5, i.e. whose 16 paths the
    GPU and the oracle follow alike (an ulp difference diverts a few per cent of the
    pixels of this glossy scene, and a diverted path's derivative is another path's)."""
    s, cams, views = c4
    spp = 16
    cfgs = [ol.settings(total_spp=spp, max_depth=8)] * len(cams)
    rng = np.random.default_rng(5)
    views.render(cfgs)
    oracles = {}
    picked = {}
    for v in (5, 21):
        s.Camera = cams[v]
        oracles[v] = ol.OracleScene(s)
        px = subset(rng, 512, 512, 2048)
        img_o, _, _ = oracles[v].render(cfgs[v], 0, spp, pixels=px)
        ref = img_o.reshape(-1, 3)[px]
        got = views.read_image(v).reshape(-1, 3)[px]
        same = np.all(np.isclose(got, ref, rtol=1e-5, atol=1e-7), axis=-1)
        print(f"C4 view {v}: {same.sum()} of {px.size} pixels follow the oracle's paths")
        assert same.mean() >= 0.75
        picked[v] = px[same]
    weights = {}
    dl = []
    for v in range(len(cams)):
        g = np.zeros((512, 512, 3), np.float32)
        if v in picked:
            w = rng.uniform(0.5, 1.5, (picked[v].size, 3)).astype(np.float32)
            g.reshape(-1, 3)[picked[v]] = w
            weights[v] = w
        dl.append(g)
    views.clear_gradients()
    views.render_adjoint(cfgs, dl, first_iteration=0)
    ground = s.Meshes[-5].Material.MaterialParameters
    for tex in (ground.BaseColor, ground.Roughness):
        idx = views.texture_index(tex)
        grad = views.gradient(idx).astype(np.float64)
        direction = rng.normal(size=grad.shape).astype(np.float32)
        ref, scale = 0.0, 0.0
        for v, px in picked.items():
            tan, nans = oracles[v].render_tangent(cfgs[v], 0, spp, idx, direction, eps=1e-3, pixels=px)
            assert nans == 0
            contrib = weights[v].astype(np.float64) * tan.reshape(-1, 3)[px]
            ref += float(contrib.sum())
            scale += float(np.abs(contrib).sum())
        got = float(np.sum(grad * direction))
        print(f"C4 adjoint texture {idx}: gpu {got:.6e} oracle {ref:.6e} (scale {scale:.3e})")
        assert scale > 0
        assert abs(got - ref) <= 0.01 * scale, (idx, got, ref, scale)


# ---------------------------------------------------------------------- C3

RES, TEX, SPP, DEPTH = 512, 256, 32, 6


@pytest.fixture(scope="module")
def c3(device):
    target_scene = scenes.box_scene(RES, RES, textured=True, tex_res=TEX, spec_tint=0.0)
    tg = GpuScene(device, target_scene)
    tg.clear()
    tg.render(ol.settings(total_spp=256, max_depth=DEPTH), 0, 256)
    target = tg.read_image().astype(np.float64)
    tg.close()
    s = scenes.box_scene(RES, RES, textured=True, tex_res=TEX, spec_tint=0.0, uniform_textures=0.5)
    return s, GpuScene(device, s), target


def test_c3_depth6_central_differences_in_expectation(c3):
    """The L2 loss linearised at the expected image, L(theta) = <w, image(theta)> with
    w = 2 (E[image] - target) / N from an independent 256-spp render at the initial
    textures: its gradient is the L2 loss's gradient at the expected image. (The L2
    loss of the very paths the adjoint replays has, in expectation, an extra
    d Var(image) / d theta term: beyond depth 4 the variance depends on the albedo
    through the Russian-roulette probability, which the detached adjoint holds fixed.)
    Per seed the same-path difference of <w, image> differs from the adjoint only by
    the Russian-roulette terms (the 1/q weights and the decisions that flip), which
    average to zero."""
    s, g, target = c3
    cfg = ol.settings(total_spp=SPP, max_depth=DEPTH)
    npx = RES * RES * 3
    mean_img = g.render_mean(ol.settings(total_spp=256, max_depth=DEPTH), 7000, 256).astype(np.float64)
    w = (2.0 * (mean_img - target) / npx).astype(np.float32)
    w64 = w.astype(np.float64)
    albedo = [g.texture_index(m.MaterialParameters.BaseColor) for m in (s.Meshes[0].Material, s.Meshes[2].Material)]
    seeds = [2000 + SPP * k for k in range(24)]  # decorrelated iteration ranges

    def loss(first):
        return float(np.sum(g.render_mean(cfg, first, SPP).astype(np.float64) * w64))

    # texels the loss depends on, from the first seed's gradient
    g.clear_gradients()
    g.render_adjoint(cfg, w, seeds[0], SPP)
    rng = np.random.default_rng(4)
    picks = []
    for idx in albedo:
        a = np.abs(g.gradient(idx))
        strong = np.argwhere(a >= 0.05 * a.max())
        picks += [(idx, tuple(strong[i])) for i in rng.choice(len(strong), 32, replace=False)]
    h = 1e-3
    fd = np.zeros((len(seeds), len(picks)))
    adj = np.zeros_like(fd)
    for k, first in enumerate(seeds):
        g.clear_gradients()
        g.render_adjoint(cfg, w, first, SPP)
        grads = {idx: g.gradient(idx).astype(np.float64) for idx in albedo}
        for j, (idx, (y, x, c)) in enumerate(picks):
            adj[k, j] = grads[idx][y, x, c]
            base = s.DeviceTextures[idx].Image
            vals = []
            for sign in (1.0, -1.0):
                t = base.copy()
                t[y, x, c] += sign * h
                g.write_texture(idx, t)
                vals.append(loss(first))
            g.write_texture(idx, base)
            fd[k, j] = (vals[0] - vals[1]) / (2 * h)
    d = fd - adj
    mean_fd, mean_adj = fd.mean(0), adj.mean(0)
    se = d.std(0, ddof=1) / np.sqrt(len(seeds))
    rel = np.abs(mean_fd - mean_adj) / np.abs(mean_adj)
    print(f"C3 depth 6: FD vs adjoint over {len(seeds)} seeds: median rel {np.median(rel):.2e}, "
          f"max rel {rel.max():.2e}, max |diff|/SE {np.max(np.abs(mean_fd - mean_adj) / (se + 1e-30)):.2f}")
    # (1) the north star's 2 % FD bar on all 64 texels together: the summed seed
    # averages of FD and adjoint, within 4 of their standard errors as well; (2) each
    # texel within 5 standard errors or 2 %. A same-path difference sees the 1/q
    # weights of Russian roulette change continuously on every seed but the survival
    # decisions that balance them in expectation flip only rarely (at h = 1e-3 a few
    # times in many seeds), so one texel's seed average carries a per-texel offset of
    # about 1 % whose sign varies from texel to texel: it shrinks as h grows (more flips
    # sampled), which a wrong derivative would not do, and cancels in the sum. The
    # exact per-texel 2 % FD check runs at MaxDepth 4, below the roulette
    # (test_gpu_adjoint.py).
    total_d = d.sum(1)
    agg = abs(float(total_d.mean()))
    agg_se = float(total_d.std(ddof=1) / np.sqrt(len(seeds)))
    agg_scale = float(np.abs(mean_adj).sum())
    print(f"C3 depth 6: summed over texels |FD - adjoint| {agg:.3e} = {agg / agg_scale:.2e} of the summed "
          f"|adjoint|, {agg / agg_se:.2f} SE")
    assert agg <= 0.02 * agg_scale and agg <= 4 * agg_se + 1e-3 * agg_scale
    z = np.abs(mean_fd - mean_adj) / (se + 1e-30)
    ok = (z <= 5) | (rel <= 0.02)
    assert ok.all(), [(picks[j], mean_fd[j], mean_adj[j], se[j]) for j in np.flatnonzero(~ok)]


def test_c3_adjoint_matches_4096spp_cpu_reference_gradient(c3):
    """The quantity: <w, dE[image]/dtheta . v> with w = dLoss/dImage of the initial
    textures (from a 256-spp render) on 512 seeded pixels, zero elsewhere. CPU: the
    oracle's tangent at 4096 spp on those pixels (8 chunks of 512 iterations give its
    standard error). GPU: the adjoint of 32 decorrelated 32-spp passes."""
    s, g, target = c3
    cfg = ol.settings(total_spp=SPP, max_depth=DEPTH)
    npx = RES * RES * 3
    img = g.render_mean(ol.settings(total_spp=256, max_depth=DEPTH), 5000, 256).astype(np.float64)
    rng = np.random.default_rng(8)
    px = subset(rng, RES, RES, 512)
    w = np.zeros((RES, RES, 3), np.float32)
    w.reshape(-1, 3)[px] = (2.0 * (img - target) / npx).reshape(-1, 3)[px]
    mats = (s.Meshes[0].Material.MaterialParameters, s.Meshes[2].Material.MaterialParameters)
    textures = [g.texture_index(m.BaseColor) for m in mats] + [g.texture_index(m.Roughness) for m in mats]
    passes = 32
    grads = {idx: [] for idx in textures}
    for k in range(passes):
        g.clear_gradients()
        g.render_adjoint(cfg, w, 10000 + SPP * k, SPP)
        for idx in textures:
            grads[idx].append(g.gradient(idx).astype(np.float64))
    o = ol.OracleScene(s)
    chunks, per = 8, 512
    for idx in textures:
        v = rng.normal(size=grads[idx][0].shape).astype(np.float32)
        gpu = np.array([float(np.sum(gk * v)) for gk in grads[idx]])
        cpu = []
        for c in range(chunks):
            c_cfg = ol.settings(total_spp=per, max_depth=DEPTH)
            tan, nans = o.render_tangent(c_cfg, 100000 + per * c, per, idx, v, eps=1e-3, pixels=px)
            assert nans == 0
            cpu.append(float(np.sum(w.reshape(-1, 3)[px].astype(np.float64) * tan.reshape(-1, 3)[px])))
        cpu = np.array(cpu)
        se = np.sqrt(gpu.var(ddof=1) / passes + cpu.var(ddof=1) / chunks)
        print(f"C3 reference gradient, texture {idx}: GPU {gpu.mean():.6e} (32 x 32 spp), "
              f"CPU {cpu.mean():.6e} (4096 spp), diff/SE {abs(gpu.mean() - cpu.mean()) / se:.2f}")
        assert abs(gpu.mean() - cpu.mean()) <= 4 * se, (idx, gpu.mean(), cpu.mean(), se)
