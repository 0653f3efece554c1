"""GPU parity tests: libsspt.so (through the C ABI) against the CPU oracle.

Bars:
* camera rays and RNG jitter: bit-identical (fp32, same operation order);
* closest-hit triangle ids and distances: identical to the oracle's brute force
  (watertight test, ties to the lower triangle id) on every ray;
* images: the GPU and the oracle draw the same RNG streams, so per-pixel
  estimates agree except where the hardware square root and reciprocal of the
  shading arithmetic (a few ulps) send a path another way. relMSE =
  mean((x - r)^2 / (r^2 + 1e-2)); bars at about 10x the observed values: the
  36-triangle box (observed 1.9e-14) <= 1e-10; the large scene at 192^2 and
  4 spp (3.3e-5: its glossy GenericMaterials amplify ulp differences into
  diverted paths) <= 3e-4 with at most 3 % of the pixels off by more than rtol
  1e-3; the
  4096-spp converged subset (1.0e-6) <= 1e-5; image means within 1 % (the
  reference's own validation bar, SeeSharp.Validation/Validator.cs:24-27).
"""
import numpy as np
import pytest

import oracle_lib as ol
from seesharp_amd import GpuScene, scenes

pytestmark = pytest.mark.gpu

SEED = 0xC030114


def relmse(x, r):
    return float(np.mean((x - r) ** 2 / (r ** 2 + 1e-2)))


@pytest.fixture(scope="module")
def box(device):
    s = scenes.box_scene(128, 128)
    return s, GpuScene(device, s), ol.OracleScene(s)


@pytest.fixture(scope="module")
def large(device):
    s = scenes.large_scene(192, 192, num_spheres=400, grid=512)
    return s, GpuScene(device, s), ol.OracleScene(s)


def test_box_primary_rays_bit_identical(box):
    _, g, o = box
    for it in (0, 7):
        rg, _ = g.primary_hits(SEED, it)
        ro, _ = o.primary_hits(SEED, it)
        assert np.array_equal(rg.view(np.uint32), ro.view(np.uint32))


def test_box_primary_hits_match_brute_force(box):
    _, g, o = box
    _, hg = g.primary_hits(SEED, 3)
    _, ho = o.primary_hits(SEED, 3, brute=True)
    assert np.array_equal(hg["prim"], ho["prim"])
    hit = ho["prim"] >= 0
    assert np.array_equal(hg["t"][hit].view(np.uint32), ho["t"][hit].view(np.uint32))


def random_rays(rng, n, lo, hi, tmax=np.inf):
    o = rng.uniform(lo, hi, size=(n, 3)).astype(np.float32)
    d = rng.normal(size=(n, 3)).astype(np.float32)
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    r = np.zeros((n, 8), np.float32)
    r[:, 0:3] = o
    r[:, 3] = 0.0
    r[:, 4:7] = d
    r[:, 7] = tmax
    return r


def test_box_random_rays_closest_and_occluded(box):
    _, g, o = box
    rays = random_rays(np.random.default_rng(1), 20000, -0.2, 1.2)
    hg = g.trace_closest(rays)
    ho = o.trace_closest(rays, brute=True)
    bad = np.flatnonzero(hg["prim"] != ho["prim"])
    assert bad.size == 0, f"{bad.size} rays differ, first {bad[:5]}: gpu {hg[bad[:5]]} oracle {ho[bad[:5]]}"
    hit = ho["prim"] >= 0
    assert np.array_equal(hg["t"][hit].view(np.uint32), ho["t"][hit].view(np.uint32))
    rays[:, 7] = 0.3
    assert np.array_equal(g.trace_occluded(rays), o.trace_occluded(rays, brute=True))


def test_box_render_matches_oracle(box):
    _, g, o = box
    cfg = ol.settings(total_spp=16, max_depth=5)
    g.clear()
    sg = g.render(cfg, 0, 16)
    img_g = g.read_image()
    img_o, so, nan_o = o.render(cfg, 0, 16)
    print(f"box 128^2 16 spp: relMSE {relmse(img_g, img_o):.3e}")
    assert g.nan_count() == nan_o == 0
    assert abs(img_g.mean() - img_o.mean()) <= 0.01 * img_o.mean()
    assert relmse(img_g, img_o) <= 1e-10
    # the ray counters follow the same paths
    assert abs(sg["num_rays"] - so.num_rays) <= 0.01 * so.num_rays
    assert abs(sg["num_shadow_rays"] - so.num_shadow_rays) <= 0.01 * so.num_shadow_rays


def test_box_render_iterations_split_equals_single_call(box):
    """FrameBuffer running mean: 4 + 4 iterations == 8 iterations in one call."""
    _, g, _ = box
    cfg = ol.settings(total_spp=8, max_depth=5)
    g.clear()
    g.render(cfg, 0, 8)
    a = g.read_image()
    g.clear()
    g.render(cfg, 0, 4)
    g.render(cfg, 4, 4)
    b = g.read_image()
    assert np.array_equal(a, b)


def test_large_primary_hits_match_oracle(large):
    _, g, o = large
    _, hg = g.primary_hits(SEED, 0)
    _, ho = o.primary_hits(SEED, 0)
    mism = np.count_nonzero(hg["prim"] != ho["prim"])
    assert mism == 0, f"{mism} of {hg.size} primary hits differ"


def test_large_random_rays_match_oracle(large):
    _, g, o = large
    rays = random_rays(np.random.default_rng(2), 50000, [-5, 0, -5], [5, 4, 5])
    hg = g.trace_closest(rays)
    ho = o.trace_closest(rays)
    bad = np.flatnonzero(hg["prim"] != ho["prim"])
    assert bad.size == 0, f"{bad.size} rays differ, first {bad[:5]}: gpu {hg[bad[:5]]} oracle {ho[bad[:5]]}"
    rays[:, 7] = 2.0
    assert np.array_equal(g.trace_occluded(rays), o.trace_occluded(rays))


def test_large_positive_octant_and_axis_parallel_rays(large):
    """Regression: directions with all components positive (an absent BVH slot's
    box [+inf,+inf]^3 passes the slab test while tmax = +inf) and directions with
    exact zero components (entry distance +inf) must traverse correctly."""
    _, g, o = large
    rng = np.random.default_rng(3)
    n = 4000
    rays = random_rays(rng, n, [-5, 0, -5], [5, 4, 5])
    rays[:, 4:7] = np.abs(rays[:, 4:7])
    axes = np.array([[1, 0, 0], [0, 1, 0], [0, 0, 1], [-1, 0, 0], [0, -1, 0], [0, 0, -1], [0.6, 0.8, 0],
                     [0, 0.6, 0.8]], np.float32)
    rays[: n // 2, 4:7] = axes[np.arange(n // 2) % len(axes)]
    hg = g.trace_closest(rays)
    ho = o.trace_closest(rays)
    bad = np.flatnonzero(hg["prim"] != ho["prim"])
    assert bad.size == 0, f"{bad.size} rays differ"
    assert np.array_equal(g.trace_occluded(rays), o.trace_occluded(rays))


def test_box_degenerate_and_extreme_directions(box):
    """The slab test's margins scale with the direction: directions shorter than
    2^-60 of the scene's scale (and zero or non-finite ones) enter every box,
    components below 2^-40 of the largest are replaced by +-2^-40 of it, and
    very long directions make tiny distances. The closest hit and occlusion
    must still equal brute force's."""
    _, g, o = box
    rng = np.random.default_rng(5)
    base = random_rays(rng, 3000, -0.2, 1.2)
    groups = []
    for scale in (1e-25, 1e-12, 1e12, 1e20):
        r = base.copy()
        r[:, 4:7] *= np.float32(scale)
        groups.append(r)
    r = base[:500].copy()
    r[:, 4:7] = 0.0  # zero direction: no hit
    groups.append(r)
    r = base[:500].copy()
    r[:, 5] = np.float32(1e-30)  # one component far below the others (and subnormal-range products)
    r[:, 6] *= np.float32(1e-41)
    groups.append(r)
    rays = np.concatenate(groups)
    hg = g.trace_closest(rays)
    ho = o.trace_closest(rays, brute=True)
    bad = np.flatnonzero(hg["prim"] != ho["prim"])
    assert bad.size == 0, f"{bad.size} rays differ, first {bad[:5]}: gpu {hg[bad[:5]]} oracle {ho[bad[:5]]}"
    hit = ho["prim"] >= 0
    assert hit.sum() > 1000
    assert np.array_equal(hg["t"][hit].view(np.uint32), ho["t"][hit].view(np.uint32))
    assert np.all(hg["prim"][3000 * 4:3000 * 4 + 500] < 0)
    assert np.array_equal(g.trace_occluded(rays), o.trace_occluded(rays, brute=True))


def test_large_render_matches_oracle(large):
    _, g, o = large
    cfg = ol.settings(total_spp=4, max_depth=8)
    g.clear()
    g.render(cfg, 0, 4)
    img_g = g.read_image()
    img_o, _, _ = o.render(cfg, 0, 4)
    div = float(np.mean(np.any(~np.isclose(img_g, img_o, rtol=1e-3, atol=1e-5), axis=-1)))
    print(f"large 192^2 4 spp: relMSE {relmse(img_g, img_o):.3e}, diverted pixels {div:.3e}")
    assert abs(img_g.mean() - img_o.mean()) <= 0.01 * img_o.mean()
    assert relmse(img_g, img_o) <= 3e-4
    assert div <= 0.03


def test_large_converged_relmse_against_cpu_reference(device):
    """North-star bar on the 1M-triangle 1024^2 depth-8 scene at 4096 spp, on a
    seeded 8192-pixel subset rendered by the CPU oracle
    (tests/golden/make_converged_reference.py):

    * same seed as the GPU (0xC030114): the GPU follows the CPU path for path,
      relMSE <= 1e-3 against the 4096-spp CPU reference (the north star's bar,
      the primary gate), and <= 1e-5, 10x the observed 1.0e-6;
    * independent seed (571298512): no error beyond Monte Carlo noise. Two
      independent 4096-spp CPU images of these pixels differ by relMSE 6.7e-3
      (2.6e-3 with 0.1 % of the pixels rejected: fireflies), so at this sample
      count an independent-seed relMSE below 1e-3 is out of reach for any
      renderer; instead the GPU's relMSE against the reference may exceed that
      CPU-vs-CPU floor by at most 25 %, both plain and with
      metrics.RelMSE_OutlierRejection (0.1 % rejected);
    * the image mean lies within 3 standard errors of the reference's.
    relMSE = mean((x - r)^2 / (r^2 + 1e-2)) per pixel and channel."""
    import os

    from seesharp_amd import metrics

    ref = np.load(os.path.join(os.path.dirname(__file__), "golden", "large_scene_ref_4096spp.npz"))
    s = scenes.large_scene(1024, 1024)
    g = GpuScene(device, s)
    spp = int(ref["spp"])
    g.clear()
    g.render(ol.settings(base_seed=int(ref["same_seed"]), total_spp=spp, max_depth=8), 0, spp)
    img = g.read_image().reshape(-1, 3)[ref["pixels"]].astype(np.float64)
    r = ref["mean"].astype(np.float64)
    same = ref["mean_same_seed"].astype(np.float64)
    err_same = relmse(img, same)
    err = relmse(img, r)
    floor = relmse(same, r)
    # the subset as a column image for the metric's per-pixel rejection
    col = lambda a: a.reshape(-1, 1, 3)  # noqa: E731
    err_rej = metrics.RelMSE_OutlierRejection(col(img), col(r))
    floor_rej = metrics.RelMSE_OutlierRejection(col(same), col(r))
    print(f"relMSE same seed {err_same:.3e}; independent {err:.3e} (CPU floor {floor:.3e}); "
          f"0.1%-outlier-rejected {err_rej:.3e} (CPU floor {floor_rej:.3e})")
    assert err_same <= 1e-3
    assert err_same <= 1e-5
    assert err <= 1.25 * floor
    assert err_rej <= 1.25 * floor_rej
    n = r.shape[0]
    # standard error of the image mean, channels treated as fully correlated (conservative)
    se_mean = np.sqrt(np.sum(ref["stderr"].astype(np.float64) ** 2) * 3) / (n * 3)
    diff = abs(float(img.mean()) - float(r.mean()))
    assert diff <= 3 * np.sqrt(2) * se_mean, (diff, se_mean)
