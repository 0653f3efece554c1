"""The reference's validation scenes (SeeSharp.Validation): small scenes built in
code or loaded from the reference's scene files, each with the sample count and
path depth its validator renders it at, and the validator's image comparison
(every image mean within 1 % of the first, SeeSharp.Validation/Validator.cs:10-30).

The reference validates its path tracer against its bidirectional integrators;
here the same scenes serve as regression fixtures for the GPU path tracer.
"""
from __future__ import annotations

import math
import os
from dataclasses import dataclass
from typing import Callable

import numpy as np

from . import sceneio
from .scene import (DiffuseEmitter, DiffuseMaterial, FrameBuffer, GenericMaterial, GlossyEmitter, Mesh,
                    PerspectiveCamera, Scene, create_look_at)


@dataclass
class ValidationScene:
    """ValidationSceneFactory (SeeSharp.Validation/ValidationSceneFactory.cs:3-9)."""
    Name: str
    SamplesPerPixel: int
    MaxDepth: int
    make: Callable[[int, int], Scene]
    # part of the reference's validator run list (SeeSharp.Validation/Program.cs:4-17)
    enabled: bool = True

    def MakeScene(self, width: int = 512, height: int = 512) -> Scene:
        return self.make(width, height)


def _quad(corners, normal=None):
    n = None if normal is None else [normal] * 4
    return Mesh(np.asarray(corners, np.float32), [0, 1, 2, 0, 2, 3], n)


def _plane_z(z, half=10.0, normal=None):
    return _quad([[-half, -half, z], [half, -half, z], [half, half, z], [-half, half, z]], normal)


def _emitter_quad(size, depth, normal_z, dx=0.0, dy=0.0):
    return _quad([[-size + dx, -size + dy, depth], [size + dx, -size + dy, depth], [size + dx, size + dy, depth],
                  [-size + dx, size + dy, depth]], (0.0, 0.0, normal_z))


def _black_diffuse():
    return DiffuseMaterial(DiffuseMaterial.Parameters(BaseColor=(0.0, 0.0, 0.0)))


def _finish(scene: Scene, camera_pos, fov, width, height, target=None):
    pos = np.asarray(camera_pos, np.float32)
    tgt = pos - np.array([0, 0, 1], np.float32) if target is None else np.asarray(target, np.float32)
    scene.Camera = PerspectiveCamera(create_look_at(pos, tgt, (0, 1, 0)), fov)
    scene.FrameBuffer = FrameBuffer(width, height)
    return scene


def direct_illum_transmit(width=512, height=512) -> Scene:
    """SeeSharp.Validation/Validate_DirectTransmit.cs:18-66: a rough, fully
    transmissive ground plane lit from behind."""
    s = Scene()
    g = _plane_z(-2)
    g.Material = GenericMaterial(GenericMaterial.Parameters(
        BaseColor=(0.8, 0.8, 0.8), Roughness=1.0, Anisotropic=1.0, IndexOfRefraction=1.45, Metallic=0.0,
        SpecularTransmittance=1.0, SpecularTintStrength=0.0))
    light = _emitter_quad(0.1, -4.0, 1.0)
    light.Material = _black_diffuse()
    s.Meshes += [g, light]
    s.Emitters.extend(DiffuseEmitter.MakeFromMesh(light, (1000.0,) * 3))
    return _finish(s, (0, 0, 0), 40, width, height)


def direct_illum(width=512, height=512) -> Scene:
    """SeeSharp.Validation/Validate_DirectIllum.cs:19-63."""
    s = Scene()
    g = _plane_z(-2)
    g.Material = GenericMaterial(GenericMaterial.Parameters(BaseColor=(1.0, 1.0, 1.0), Roughness=0.5))
    light = _emitter_quad(0.1, -1.0, -1.0)
    light.Material = _black_diffuse()
    s.Meshes += [g, light]
    s.Emitters.extend(DiffuseEmitter.MakeFromMesh(light, (1000.0,) * 3))
    return _finish(s, (0, 0, 0), 40, width, height)


def single_bounce(width=512, height=512) -> Scene:
    """SeeSharp.Validation/Validate_SingleBounce.cs:19-80: light reaches the
    ground only after one bounce off a reflector behind the camera."""
    s = Scene()
    g = _plane_z(-2)
    g.Material = GenericMaterial(GenericMaterial.Parameters(BaseColor=(1.0, 1.0, 1.0)))
    light = _emitter_quad(0.1, 1.0, 1.0)
    light.Material = GenericMaterial(GenericMaterial.Parameters(BaseColor=(0.0, 0.0, 0.0)))
    refl = _plane_z(2.0, normal=(0.0, 0.0, 1.0))
    refl.Material = GenericMaterial(GenericMaterial.Parameters(BaseColor=(1.0, 1.0, 1.0)))
    s.Meshes += [g, light]
    s.Emitters.extend(DiffuseEmitter.MakeFromMesh(light, (1000.0,) * 3))
    s.Meshes.append(refl)
    return _finish(s, (0, 0, 0), 40, width, height)


def single_bounce_glossy(width=512, height=512) -> Scene:
    """SeeSharp.Validation/Validate_SingleBounceGlossy.cs:19-89: two planes, a
    rough transmissive one in front of a rough reflective one, lit from far
    behind."""
    s = Scene()
    front = _quad([[-1, -1, 0], [1, -1, 0], [1, 1, 0], [-1, 1, 0]])
    front.Material = GenericMaterial(GenericMaterial.Parameters(BaseColor=(1.0, 1.0, 1.0), Roughness=0.5,
                                                                SpecularTransmittance=1.0))
    back = _quad([[-1, -1, 1], [1, -1, 1], [1, 1, 1], [-1, 1, 1]])
    back.Material = GenericMaterial(GenericMaterial.Parameters(BaseColor=(1.0, 1.0, 1.0), Roughness=0.5))
    area = np.float32(0.1)
    pos = float(np.sqrt(area) / np.float32(2))
    light = _emitter_quad(pos, 11.0, -1.0)
    light.Material = _black_diffuse()
    s.Meshes += [front, back, light]
    radiance = float(np.float32(500) / (np.float32(math.pi) * area))
    s.Emitters.extend(DiffuseEmitter.MakeFromMesh(light, (radiance,) * 3))
    return _finish(s, (0, 0, -1), 45, width, height, target=(0, 0, 0))


def multi_light(width=512, height=512) -> Scene:
    """SeeSharp.Validation/Validate_MultiLight.cs:19-66: two small lights."""
    s = Scene()
    g = _plane_z(-2)
    g.Material = DiffuseMaterial(DiffuseMaterial.Parameters(BaseColor=(1.0, 1.0, 1.0)))
    s.Meshes.append(g)
    for dx in (-1.0, 1.0):
        light = _emitter_quad(0.1, -1.0, -1.0, dx=dx)
        light.Material = _black_diffuse()
        s.Meshes.append(light)
        s.Emitters.extend(DiffuseEmitter.MakeFromMesh(light, (1000.0,) * 3))
    return _finish(s, (0, 0, 0), 40, width, height)


def glossy_light(width=512, height=512) -> Scene:
    """SeeSharp.Validation/Validate_GlossyLight.cs:19-66: a glossy emitter
    (exponent 200) behind a diffuse transmitter."""
    s = Scene()
    g = _plane_z(0)
    g.Material = DiffuseMaterial(DiffuseMaterial.Parameters(BaseColor=(1.0, 1.0, 1.0), Transmitter=True))
    light = _emitter_quad(0.1, -1.0, 1.0)
    light.Material = _black_diffuse()
    s.Meshes += [g, light]
    s.Emitters.extend(GlossyEmitter.MakeFromMesh(light, (1000.0,) * 3, 200.0))
    return _finish(s, (0, 0, 2), 40, width, height, target=(0, 0, -1))


def _from_file(rel):
    def make(width, height, scenes_dir=None):
        return sceneio.load_scene(os.path.join(scenes_dir or default_scenes_dir(), rel), width, height)
    return make


def default_scenes_dir() -> str:
    """The reference's scene files (Data/Scenes) as copied beside the tests."""
    here = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    return os.environ.get("SSPT_SCENES_DIR", os.path.join(here, "tests", "golden", "scenes"))


SCENES = [
    ValidationScene("DirectIllumTransmit", 10, 2, direct_illum_transmit),
    ValidationScene("DirectIllumination", 10, 2, direct_illum),
    ValidationScene("SingleBounce", 10, 3, single_bounce),
    # "Needs more samples to correctly validate" (SeeSharp.Validation/Program.cs:9-10)
    ValidationScene("SingleBounceGlossy", 32, 3, single_bounce_glossy, enabled=False),
    ValidationScene("MultiLight", 10, 2, multi_light),
    ValidationScene("GlossyLight", 10, 2, glossy_light),
    # SeeSharp.Validation/Validate_Environment.cs:12-16
    ValidationScene("Environment", 10, 5, _from_file("simplebackground.json")),
    # SeeSharp.Validation/Validate_CornellBox.cs:12-16
    ValidationScene("CornellBox", 8, 5, _from_file(os.path.join("CornellBox", "CornellBox.json"))),
    # SeeSharp.Validation/Validate_Textures.cs:12-16 (rendered at 700 x 500 there)
    ValidationScene("TextureTest", 8, 5, _from_file(os.path.join("TextureTest", "TextureTest.json"))),
]


def by_name(name: str) -> ValidationScene:
    for v in SCENES:
        if v.Name == name:
            return v
    raise KeyError(name)


def image_mean(img: np.ndarray) -> float:
    """The validator's per-image average of (R + G + B) / 3 (Validator.cs:12-20),
    summed in float64."""
    return float(np.asarray(img, np.float64).mean())


def means_agree(images, tolerance: float = 0.01) -> bool:
    """ValidateImages (Validator.cs:10-30): every mean within `tolerance` of the first's."""
    means = [image_mean(i) for i in images]
    return all(abs(m - means[0]) <= means[0] * tolerance for m in means)
