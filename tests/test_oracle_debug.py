"""The oracle's DebugVisualizer (SeeSharp/Integrators/DebugVisualizer.cs) on
the CPU against closed forms: a camera facing a plane sees |cos| between the
plane's normal and each jittered primary ray, i.e. |d . n| for the oracle's
own primary rays, averaged over the iterations; misses are black."""
import numpy as np

import oracle_lib as ol
from seesharp_amd import scenes
from seesharp_amd.scene import DiffuseEmitter, DiffuseMaterial, FrameBuffer, Mesh, PerspectiveCamera, Scene, create_look_at


def plane_scene(tilt):
    """A plane through the origin with normal (0, sin t, cos t); camera on +z."""
    c, s = np.cos(tilt), np.sin(tilt)
    u = np.array([1.0, 0, 0])
    v = np.array([0, c, -s])
    pts = [(-3 * u - 3 * v), (3 * u - 3 * v), (3 * u + 3 * v), (-3 * u + 3 * v)]
    m = Mesh(np.array(pts, np.float32), [0, 1, 2, 0, 2, 3])
    m.Material = DiffuseMaterial(DiffuseMaterial.Parameters())
    # a light behind the camera makes the scene valid (Scene.IsValid) and stays out of view
    light = Mesh(np.array([[0, 0, 9], [1, 0, 9], [0, 1, 9]], np.float32), [0, 1, 2])
    light.Material = DiffuseMaterial(DiffuseMaterial.Parameters(BaseColor=(0, 0, 0)))
    sc = Scene()
    sc.Meshes += [m, light]
    sc.Emitters.extend(DiffuseEmitter.MakeFromMesh(light, (1.0, 1.0, 1.0)))
    sc.Camera = PerspectiveCamera(create_look_at((0, 0, 4), (0, 0, 0), (0, 1, 0)), 60)
    sc.FrameBuffer = FrameBuffer(24, 20)
    return sc, np.array([0, s, c])


def test_debug_view_of_a_plane_closed_form():
    for tilt in (0.0, 0.5, 1.1):
        sc, n = plane_scene(tilt)
        o = ol.OracleScene(sc)
        img = o.render_debug(0xC030114, 0, 4)
        expected = np.zeros(24 * 20)
        for it in range(4):
            rays, hits = o.primary_hits(0xC030114, it)
            d = rays[:, 4:7].astype(np.float64)
            cos = np.abs(d @ n) / np.linalg.norm(d, axis=1)
            expected += np.where(np.isin(hits["prim"], (0, 1)), cos, 0.0)
        expected /= 4
        assert np.allclose(img[..., 0].reshape(-1), expected, rtol=1e-5, atol=1e-6), tilt
        assert np.array_equal(img[..., 0], img[..., 1]) and np.array_equal(img[..., 0], img[..., 2])
        o.close()


def test_debug_view_is_a_running_mean():
    s = scenes.box_scene(16, 16)
    o = ol.OracleScene(s)
    a = o.render_debug(7, 0, 6)
    # iterations 0..5 in one call equal the mean of six single-iteration images
    singles = [o.render_debug(7, it, 1) * (it + 1) for it in range(6)]  # undo the 1/(it+1) of a lone splat
    assert np.allclose(a, np.mean(singles, axis=0), rtol=1e-5, atol=1e-7)
    assert 0 < a.mean() < 1 and a.max() <= 1.0 + 1e-6
    o.close()
