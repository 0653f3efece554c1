"""GPU DebugVisualizer (sspt_render_debug) against the oracle's restatement:
the eye-light images agree to fp32 rounding on every pixel, and the
integrator class renders the reference's CornellBox.json."""
import os

import numpy as np
import pytest

import oracle_lib as ol
from seesharp_amd import DebugVisualizer, GpuScene, sceneio, scenes

pytestmark = pytest.mark.gpu

SCENES = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "scenes")


@pytest.mark.parametrize("which", ["box", "large"])
def test_debug_view_matches_oracle(device, which):
    s = scenes.box_scene(64, 48) if which == "box" else scenes.large_scene(96, 64, num_spheres=80, grid=64)
    g = GpuScene(device, s)
    o = ol.OracleScene(s)
    g.clear()
    g.render_debug(0xC030114, 0, 3)
    g.render_debug(0xC030114, 3, 2)  # running mean over calls
    img = g.read_image()
    ref = o.render_debug(0xC030114, 0, 5)
    o.close()
    assert ref.mean() > 0
    assert np.allclose(img, ref, rtol=1e-5, atol=1e-6)


def test_debug_visualizer_integrator(device):
    s = sceneio.load_scene(os.path.join(SCENES, "CornellBox", "CornellBox.json"), 64, 64)
    dv = DebugVisualizer(device)
    dv.TotalSpp = 4
    fb = dv.Render(s)
    o = ol.OracleScene(s)
    ref = o.render_debug(dv.BaseSeed, 0, 4)
    o.close()
    assert fb.CurIteration == 4 and fb.MetaData["NumIterations"] == 4
    assert np.allclose(fb.Image, ref, rtol=1e-5, atol=1e-6)
    assert 0.2 < fb.Image.mean() < 1.0
