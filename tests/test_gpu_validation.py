"""GPU renders of the reference's validation scenes (seesharp_amd.validation,
after SeeSharp.Validation) at the validator's resolution, sample count and
depth, against the oracle:

* same seed: the GPU and the oracle follow the same paths, so image means agree
  within 1 % and relMSE <= 1e-9 (observed 5e-16 .. 1.7e-11, the largest on the
  rough dielectric of DirectIllumTransmit);
* independent seeds: the validator's own bar, every image mean within 1 % of
  the first (SeeSharp.Validation/Validator.cs:10-30), GPU against oracle;
* MultiLight against Lambert's closed form on every ground pixel.
"""
import math

import numpy as np
import pytest

import oracle_lib as ol
from seesharp_amd import GpuScene, validation
from test_validation_scenes import lambert_irradiance_many

pytestmark = pytest.mark.gpu

REF_SEED = 571298512  # the reference's own seed for reference renders


def relmse(x, r):
    return float(np.mean((x - r) ** 2 / (r ** 2 + 1e-2)))


def size_of(v):
    return (700, 500) if v.Name == "TextureTest" else (512, 512)


@pytest.mark.parametrize("name", [v.Name for v in validation.SCENES if v.enabled])
def test_validation_scene_matches_oracle(device, name):
    v = validation.by_name(name)
    s = v.MakeScene(*size_of(v))
    g = GpuScene(device, s)
    o = ol.OracleScene(s)
    spp, depth = v.SamplesPerPixel, v.MaxDepth
    same = ol.settings(base_seed=REF_SEED, total_spp=spp, max_depth=depth)
    g.clear()
    g.render(same, 0, spp)
    img_g = g.read_image()
    img_o, _, nan_o = o.render(same, 0, spp)
    other, _, _ = o.render(ol.settings(base_seed=0xC030114, total_spp=spp, max_depth=depth), 0, spp)
    o.close()
    print(f"{name}: gpu {validation.image_mean(img_g):.5f} oracle {validation.image_mean(img_o):.5f} "
          f"oracle (other seed) {validation.image_mean(other):.5f} relMSE {relmse(img_g, img_o):.2e}")
    assert g.nan_count() == nan_o == 0
    assert validation.image_mean(img_o) > 0
    assert abs(validation.image_mean(img_g) - validation.image_mean(img_o)) <= 0.01 * validation.image_mean(img_o)
    assert relmse(img_g, img_o) <= 1e-9
    assert validation.means_agree([img_g, other])


def test_multi_light_closed_form_on_gpu(device):
    """The GPU image of MultiLight (256 spp, MaxDepth 2) against (1/pi) * 1000 *
    the Lambert irradiance of both lights at each iteration's ground point."""
    s = validation.by_name("MultiLight").MakeScene(32, 32)
    polys = [m.Vertices.astype(np.float64) for m in s.Meshes[1:]]
    g = GpuScene(device, s)
    spp = 256
    g.clear()
    g.render(ol.settings(total_spp=spp, max_depth=2), 0, spp)
    img = g.read_image()
    acc = np.zeros(32 * 32)
    ground = np.ones(32 * 32, bool)
    for it in range(spp):
        rays, hits = g.primary_hits(0xC030114, it)
        ground &= np.isin(hits["prim"], (0, 1))
        p = rays[:, :3].astype(np.float64) + hits["t"][:, None] * rays[:, 4:7].astype(np.float64)
        acc += sum(lambert_irradiance_many(p, np.array([0, 0, 1.0]), poly) for poly in polys)
    expected = 1000.0 * acc / spp / math.pi
    ratio = img.reshape(-1, 3).mean(axis=1)[ground] / expected[ground]
    print(f"MultiLight closed form: mean ratio {ratio.mean():.5f}, max deviation {np.abs(ratio - 1).max():.3f}")
    assert ground.sum() > 800
    assert np.all(np.abs(ratio - 1) < 0.25)
    assert abs(ratio.mean() - 1) < 0.01
