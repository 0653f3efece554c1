"""GPU FrameBuffer features through the C ABI: the outlier cache against the
oracle's restatement (the same quaternary heap fed in iteration order), in
running-mean renders split over calls and in batch renders; its link to
ReplayPixel; and the path tracer's per-iteration error log and file output."""
import json
import os

import numpy as np
import pytest

import oracle_lib as ol
from seesharp_amd import GpuBatch, GpuPathTracer, GpuScene, imageio, scenes
from seesharp_amd.scene import FrameBuffer

pytestmark = pytest.mark.gpu


def agree(gw, gi, ow, oi, max_frac=0.01):
    """Tracked samples equal (same iterations, weights to fp32 summation order)
    except in pixels where a sample of the GPU and the oracle went different
    ways (a last-ulp difference changing a sampling decision) or two near-equal
    weights swapped: at most max_frac of the pixels."""
    ok = np.all(gi == oi, axis=-1) & np.all(np.isclose(gw, ow, rtol=1e-5, atol=1e-7), axis=-1)
    bad = np.count_nonzero(~ok)
    print(f"outlier queues differing in {bad} of {ok.size} pixels")
    assert bad <= max_frac * ok.size, bad


@pytest.mark.parametrize("n,spp", [(4, 16), (2, 9), (4, 3), (1, 12)])
def test_outliers_match_oracle(device, n, spp):
    s = scenes.large_scene(64, 48, num_spheres=40, grid=32)
    g = GpuScene(device, s)
    o = ol.OracleScene(s)
    cfg = ol.settings(max_depth=6)
    g.track_outliers(n)
    g.clear()
    g.render(cfg, 0, spp)
    gw, gi = g.read_outliers()
    ow, oi = o.render_outliers(cfg, 0, spp, n)
    o.close()
    assert gw.shape == (48, 64, n)
    agree(gw, gi, ow, oi)


def test_outliers_across_calls_waves_and_batches(device):
    """Split calls, small waves (several waves per call) and a batch world (the
    slot-parallel accumulation) all feed each pixel's queue in iteration order:
    identical queues."""
    s = scenes.box_scene(40, 30)
    cfg = ol.settings(max_depth=5)
    g = GpuScene(device, s)
    g.track_outliers(3)
    g.clear()
    g.render(cfg, 0, 11)
    ref = g.read_outliers()
    g.clear()
    g.render(cfg, 0, 4)
    g.render(cfg, 4, 7)
    a = g.read_outliers()
    # waves of two iterations for the running mean; the batch's waves are cut
    # mid-iteration at this capacity
    device.set_wave_capacity(40 * 30 * 2 + 17)
    g.clear()
    g.render(cfg, 0, 11)
    b = g.read_outliers()
    other = scenes.large_scene(24, 16, num_spheres=10, grid=8)
    batch = GpuBatch(device, [other, s])
    batch.track_outliers(3)
    batch.render([ol.settings(total_spp=5, max_depth=5), ol.settings(total_spp=11, max_depth=5)])
    c = batch.read_outliers(1)
    device.set_wave_capacity(0)
    for x in (a, b, c):
        assert np.array_equal(x[0], ref[0]) and np.array_equal(x[1], ref[1])
    # tracking off: reading is an error
    g.track_outliers(0)
    with pytest.raises(RuntimeError):
        g.read_outliers()


def test_largest_outlier_replays_to_its_weight(device):
    """OutlierCacheTest's use (SeeSharp.IntegrationTests/OutlierCacheTest.cs:24-37):
    replaying the strongest sample's iteration gives that sample."""
    s = scenes.box_scene(32, 32)
    g = GpuScene(device, s)
    cfg = ol.settings(max_depth=8)
    g.track_outliers(4)
    g.clear()
    g.render(cfg, 0, 20)
    w, it = g.read_outliers()
    for col, row in ((5, 7), (16, 16), (30, 2), (11, 25)):
        _, est = g.replay_pixel(cfg, col, row, int(it[row, col, 0]))
        avg = (np.float32(est[0]) + np.float32(est[1]) + np.float32(est[2])) / np.float32(3)
        assert avg == w[row, col, 0]


def test_path_tracer_error_log_and_file(device, tmp_path):
    s = scenes.box_scene(48, 48)
    ref_pt = GpuPathTracer(device)
    ref_pt.TotalSpp = 256
    ref_pt.MaxDepth = 5
    ref_pt.BaseSeed = 571298512
    ref_pt.Render(s)
    reference = s.FrameBuffer.Image.copy()
    s.FrameBuffer = FrameBuffer(48, 48, str(tmp_path / "box.exr"))
    s.FrameBuffer.ReferenceImage = reference
    pt = GpuPathTracer(device)
    pt.TotalSpp = 16
    pt.MaxDepth = 5
    fb = pt.Render(s)
    assert len(fb.Errors) == 16
    rel = [e["RelMSE"] for e in fb.Errors]
    assert rel[-1] < rel[0] and all(e["RenderTime"] > 0 for e in fb.Errors)
    # the per-iteration render equals the one-call render
    one = GpuScene(device, s)
    one.clear()
    one.render(pt.settings(), 0, 16)
    assert np.allclose(fb.Image, one.read_image(), rtol=1e-6, atol=1e-7)
    assert fb.Outliers is not None and fb.Outliers[1].shape == (48, 48, 4)
    fb.WriteToFile()
    layers = imageio.read_exr_layers(str(tmp_path / "box.exr"))
    assert {"", "albedo", "normal", "outlier-0", "outlier-3"} <= set(layers)
    assert np.array_equal(layers[""], fb.Image)
    meta = json.load(open(tmp_path / "box.json"))
    assert len(meta["ErrorMetrics"]) == 16 and meta["NumIterations"] == 16
    assert os.path.getsize(tmp_path / "box.exr") > 0
