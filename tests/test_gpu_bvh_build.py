"""Device BVH build (sspt_set_bvh_builder, SSPT_BVH_DEVICE) against the host
builder, the CommitScene step of Scene.Prepare (SeeSharp/Scene.cs:108-113).

Both builders run the same binned SAH with the same fp32 split costs and stable
partitions, so the trees they make must agree node for node: equal child boxes,
child indices, counts and slot masks, and the same triangle at every leaf
position, also where a node's triangle centroids all coincide and it takes the
object-median split (duplicated triangles, the two triangles of a quad). On top
of the structural check, renders through the two trees are bit-identical, and
the 1 036 296-triangle large scene builds in at most 50 ms of kernel time."""
import numpy as np
import pytest

from seesharp_amd import GpuScene, scenes
from seesharp_amd.scene import (DiffuseEmitter, DiffuseMaterial, FrameBuffer, Mesh, PerspectiveCamera, Scene,
                                create_look_at)

import oracle_lib as ol

pytestmark = pytest.mark.gpu


def soup_scene(tris: np.ndarray) -> Scene:
    """A scene of the (n, 3, 3) triangles plus one emissive triangle above them."""
    s = Scene()
    if tris.shape[0]:
        m = Mesh(tris.reshape(-1, 3), np.arange(tris.shape[0] * 3, dtype=np.int32))
        m.Material = DiffuseMaterial(DiffuseMaterial.Parameters(BaseColor=(0.7, 0.7, 0.7)))
        s.Meshes.append(m)
    light = Mesh(np.array([[-1, 3, -1], [1, 3, -1], [0, 3, 1]], np.float32), np.array([0, 2, 1], np.int32))
    light.Material = DiffuseMaterial(DiffuseMaterial.Parameters(BaseColor=(0.0, 0.0, 0.0)))
    s.Meshes.append(light)
    s.Emitters.extend(DiffuseEmitter.MakeFromMesh(light, (5.0, 5.0, 5.0)))
    s.Camera = PerspectiveCamera(create_look_at((0.0, 1.0, 4.0), (0.0, 0.0, 0.0), (0, 1, 0)), 50.0)
    s.FrameBuffer = FrameBuffer(32, 32)
    return s


def random_soup(n: int, seed: int, flat_axis: int | None = None) -> np.ndarray:
    rng = np.random.default_rng(seed)
    centre = rng.uniform(-2.0, 2.0, (n, 1, 3))
    tris = (centre + rng.normal(0.0, 0.05, (n, 3, 3))).astype(np.float32)
    if flat_axis is not None:
        tris[:, :, flat_axis] = 0.25
    return tris


def build_both(device, scene):
    device.set_bvh_builder(host=True)
    try:
        host = GpuScene(device, scene)
    finally:
        device.set_bvh_builder(host=False)
    dev = GpuScene(device, scene)
    assert dev.bvh_build_time()["builder"] == "device"
    assert host.bvh_build_time()["builder"] == "host"
    return dev, host


def assert_same_tree(dev, host):
    nd, td = dev.read_bvh()
    nh, th = host.read_bvh()
    assert nd.shape == nh.shape, f"{nd.shape[0]} device nodes vs {nh.shape[0]} host nodes"
    # box planes: equal as floats (a bound may be -0 on one side and +0 on the other)
    planes_equal = nd[:, :24] == nh[:, :24]
    assert planes_equal.all(), f"{int((~planes_equal).sum())} box planes differ"
    ints_d, ints_h = nd[:, 24:].view(np.int32), nh[:, 24:].view(np.int32)
    assert np.array_equal(ints_d, ints_h), f"{int((ints_d != ints_h).sum())} child index / count words differ"
    assert np.array_equal(td, th), f"{int((td != th).sum())} leaf positions hold other triangles"
    return nd.shape[0]


@pytest.mark.parametrize("n", [0, 1, 2, 3, 5, 64, 255, 256, 257, 2047, 2049, 5000])
def test_small_soups_match_host(device, n):
    # n = 0: the emissive triangle alone, a root that is a leaf
    s = soup_scene(random_soup(n, seed=n))
    dev, host = build_both(device, s)
    nodes = assert_same_tree(dev, host)
    print(f"n={n}: {nodes} nodes, device kernels {dev.bvh_build_time()['kernel_ms']:.3f} ms")


@pytest.mark.parametrize("flat_axis", [0, 1, 2])
def test_planar_soup_matches_host(device, flat_axis):
    # every centroid has the same coordinate on one axis: that axis never splits
    s = soup_scene(random_soup(20000, seed=7 + flat_axis, flat_axis=flat_axis))
    dev, host = build_both(device, s)
    assert_same_tree(dev, host)


def test_large_soup_matches_host(device):
    s = soup_scene(random_soup(300000, seed=11))
    dev, host = build_both(device, s)
    nodes = assert_same_tree(dev, host)
    t = dev.bvh_build_time()
    print(f"300k soup: {nodes} nodes, device kernels {t['kernel_ms']:.2f} ms, wall {t['wall_ms']:.1f} ms, "
          f"host wall {host.bvh_build_time()['wall_ms']:.1f} ms")


def test_duplicated_triangles_match_host(device):
    # 4 copies of each of 3000 triangles: every group of copies ends in an object-median split
    base = random_soup(3000, seed=3)
    s = soup_scene(np.repeat(base, 4, axis=0))
    dev, host = build_both(device, s)
    assert_same_tree(dev, host)


def test_box_renders_identically(device):
    s = scenes.box_scene(64, 64)
    dev, host = build_both(device, s)
    assert_same_tree(dev, host)
    cfg = ol.settings(base_seed=0xC030114, total_spp=8, max_depth=5)
    images = []
    for g in (dev, host):
        g.clear()
        g.render(cfg, 0, 8)
        images.append(g.read_image())
    assert images[0].mean() > 0
    assert np.array_equal(images[0], images[1])


def test_large_scene_device_build(device):
    s = scenes.large_scene(128, 128)
    dev, host = build_both(device, s)
    nodes = assert_same_tree(dev, host)
    t, th = dev.bvh_build_time(), host.bvh_build_time()
    print(f"large scene: {dev.info()['num_triangles']} triangles, {nodes} nodes; device kernels "
          f"{t['kernel_ms']:.2f} ms (wall {t['wall_ms']:.1f} ms), host wall {th['wall_ms']:.1f} ms")
    assert t["kernel_ms"] <= 50.0
    cfg = ol.settings(base_seed=0xC030114, total_spp=2, max_depth=5)
    images = []
    for g in (dev, host):
        g.clear()
        g.render(cfg, 0, 2)
        images.append(g.read_image())
    assert np.array_equal(images[0], images[1])
