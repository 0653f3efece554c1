"""BVH leaf sizes (sspt_set_bvh_leaf_size) through the traversal.

Single-triangle leaves queue one triangle per FIFO entry; larger leaf sizes
(the default is two) make the traversal nodes' other leaf kinds: leaves of two
to four triangles, queued as one entry, and leaves of more than four, read
from the big-leaf table and tested in place. The closest hit is decided by distance
and then by the lower triangle id whatever the tree, so every leaf size must
give brute force's closest hits and occlusion answers, and renders identical
bit for bit to the single-triangle tree's, with the same ray counts."""
import numpy as np
import pytest

import oracle_lib as ol
from seesharp_amd import GpuScene, scenes
from test_gpu_bvh_build import random_soup, soup_scene
from test_gpu_parity import random_rays

pytestmark = pytest.mark.gpu

SIZES = (2, 4, 8, 16)


def build(device, scene, leaf):
    before = device.bvh_leaf_size()
    device.set_bvh_leaf_size(leaf)
    try:
        return GpuScene(device, scene)
    finally:
        device.set_bvh_leaf_size(before)


def leaf_counts(g):
    nodes, _ = g.read_bvh()
    c = nodes[:, 28:32].view(np.int32).copy()
    c[:, 0] = (c[:, 0] << 8) >> 8
    return c[c > 0]


@pytest.fixture(scope="module")
def soup():
    s = soup_scene(random_soup(3000, 21))
    o = ol.OracleScene(s)
    yield s, o
    o.close()


@pytest.mark.parametrize("leaf", SIZES)
def test_soup_rays_match_brute_force(device, soup, leaf):
    s, o = soup
    g = build(device, s, leaf)
    counts = leaf_counts(g)
    assert counts.max() <= leaf and counts.max() > 1
    if leaf > 4:
        assert np.any(counts > 4), "no leaf of more than four triangles"
    rays = random_rays(np.random.default_rng(leaf), 20000, -2.5, 2.5)
    hg = g.trace_closest(rays)
    ho = o.trace_closest(rays, brute=True)
    bad = np.flatnonzero(hg["prim"] != ho["prim"])
    assert bad.size == 0, f"leaf {leaf}: {bad.size} rays differ, first {bad[:5]}"
    hit = ho["prim"] >= 0
    assert hit.mean() > 0.05
    assert np.array_equal(hg["t"][hit].view(np.uint32), ho["t"][hit].view(np.uint32))
    rays[:, 7] = 0.5
    assert np.array_equal(g.trace_occluded(rays), o.trace_occluded(rays, brute=True))


def test_large_scene_renders_equal_single_triangle_leaves(device):
    # leaf size 1 is the reference tree here; SIZES include the default 2
    s = scenes.large_scene(96, 96, num_spheres=60, grid=64)
    cfg = ol.settings(total_spp=4, max_depth=8)
    base = build(device, s, 1)
    base.clear()
    st0 = base.render(cfg, 0, 4)
    img0 = base.read_image()
    for leaf in SIZES:
        g = build(device, s, leaf)
        assert leaf_counts(g).max() > 1
        g.clear()
        st = g.render(cfg, 0, 4)
        img = g.read_image()
        assert np.array_equal(img.view(np.uint32), img0.view(np.uint32)), f"leaf {leaf}: image differs"
        for k in ("num_rays", "num_shadow_rays", "num_occluded", "num_ray_hits"):
            assert st[k] == st0[k], (leaf, k)


def test_leaf_size_is_checked(device):
    from seesharp_amd import abi

    before = device.bvh_leaf_size()
    assert before == 2  # the library's default
    for bad in (0, 17, -1):
        with pytest.raises(abi.SsptError):
            device.set_bvh_leaf_size(bad)
    assert device.bvh_leaf_size() == before
