"""Host-side mirror of SeeSharp's scene object model.

The classes keep the reference's names, constructor arguments and defaults so
that code written against SeeSharp reads the same here:

* ``Mesh`` -- SeeSharp/Geometry/Mesh.cs:6-30
* ``TextureRgb`` / ``TextureMono`` -- SeeSharp/Images/TextureRgb.cs, TextureMono.cs
* ``DiffuseMaterial`` -- SeeSharp/Shading/Materials/DiffuseMaterial.cs:6-20,193
* ``GenericMaterial`` -- SeeSharp/Shading/Materials/GenericMaterial.cs:3-60
* ``DiffuseEmitter`` -- SeeSharp/Shading/Emitters/DiffuseEmitter.cs:6-30
* ``EnvironmentMap`` -- SeeSharp/Shading/Background/EnvironmentMap.cs:13-17
* ``PerspectiveCamera`` -- SeeSharp/Cameras/PerspectiveCamera.cs:29-34
* ``FrameBuffer`` -- SeeSharp/Images/FrameBuffer.cs:78-160
* ``Scene`` -- SeeSharp/Scene.cs:6-164

``Scene.to_desc()`` flattens the scene into the ``sspt_scene_desc`` of
include/sspt.h in Meshes / Emitters order; both libsspt.so and the test
oracle consume that description.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import abi


def f32(a, shape_last=None):
    arr = np.ascontiguousarray(np.asarray(a, dtype=np.float32))
    if shape_last is not None:
        arr = arr.reshape(-1, shape_last)
    return arr


class Mesh:
    """SeeSharp/Geometry/Mesh.cs:6-30 (positions, indices, optional shading normals and uvs)."""

    def __init__(self, vertices, indices, shadingNormals=None, textureCoordinates=None):
        self.Vertices = f32(vertices, 3)
        self.Indices = np.ascontiguousarray(np.asarray(indices, dtype=np.int32).reshape(-1))
        if self.Indices.size % 3:
            raise ValueError("index count must be a multiple of 3")
        n = self.Vertices.shape[0]
        if self.Indices.size and (self.Indices.min() < 0 or self.Indices.max() >= n):
            raise ValueError("index out of range")
        self.ShadingNormals = None if shadingNormals is None else f32(shadingNormals, 3)
        self.TextureCoordinates = None if textureCoordinates is None else f32(textureCoordinates, 2)
        self.Material = None
        self.Name = ""

    @property
    def NumFaces(self) -> int:
        return self.Indices.size // 3

    @property
    def FaceNormals(self) -> np.ndarray:
        v = self.Vertices.astype(np.float32)
        i = self.Indices.reshape(-1, 3)
        n = np.cross(v[i[:, 1]] - v[i[:, 0]], v[i[:, 2]] - v[i[:, 0]])
        return (n / np.linalg.norm(n, axis=1, keepdims=True)).astype(np.float32)


class TextureRgb:
    """Constant color or RGB image with nearest-texel lookup (ImageTexture.cs:27-44)."""

    Repeat, Clamp = 0, 1

    def __init__(self, value):
        arr = np.asarray(value, dtype=np.float32)
        if arr.ndim == 3:
            self.Image = np.ascontiguousarray(arr[..., :3])
            self.constColor = (0.0, 0.0, 0.0)
        else:
            self.Image = None
            self.constColor = tuple(float(x) for x in arr.reshape(3))
        self.Border = TextureRgb.Repeat

    @property
    def IsConstant(self):
        return self.Image is None


class TextureMono:
    """Constant value or monochrome image (TextureMono.cs:6-38)."""

    Repeat, Clamp = 0, 1

    def __init__(self, value):
        arr = np.asarray(value, dtype=np.float32)
        if arr.ndim == 2:
            self.Image = np.ascontiguousarray(arr)
            self.constColor = 0.0
        else:
            self.Image = None
            self.constColor = float(arr)
        self.Border = TextureMono.Repeat

    @property
    def IsConstant(self):
        return self.Image is None


def _tex_rgb(v):
    return v if isinstance(v, TextureRgb) else TextureRgb(v)


def _tex_mono(v):
    return v if isinstance(v, TextureMono) else TextureMono(v)


class DiffuseMaterial:
    """SeeSharp/Shading/Materials/DiffuseMaterial.cs:6-20."""

    @dataclass
    class Parameters:
        BaseColor: object = (1.0, 1.0, 1.0)
        Transmitter: bool = False

    def __init__(self, parameters: "DiffuseMaterial.Parameters"):
        self.MaterialParameters = parameters
        self.Name = ""


class GenericMaterial:
    """SeeSharp/Shading/Materials/GenericMaterial.cs:3-60. Class defaults:
    Roughness 0.5, SpecularTintStrength 1.0, IndexOfRefraction 1.45."""

    @dataclass
    class Parameters:
        BaseColor: object = (1.0, 1.0, 1.0)
        Roughness: object = 0.5
        Metallic: float = 0.0
        SpecularTintStrength: float = 1.0
        Anisotropic: float = 0.0
        SpecularTransmittance: float = 0.0
        IndexOfRefraction: float = 1.45

    def __init__(self, parameters: "GenericMaterial.Parameters"):
        self.MaterialParameters = parameters
        self.Name = ""


class DiffuseEmitter:
    """One emissive triangle (SeeSharp/Shading/Emitters/DiffuseEmitter.cs:6-30)."""

    def __init__(self, mesh: Mesh, face: int, radiance):
        self.Mesh = mesh
        self.FaceIndex = int(face)
        self.Radiance = tuple(float(x) for x in np.asarray(radiance, dtype=np.float32).reshape(3))

    @staticmethod
    def MakeFromMesh(mesh: Mesh, radiance):
        """DiffuseEmitter.cs:17-30: one emitter per face, zero-area faces skipped."""
        v = mesh.Vertices
        i = mesh.Indices.reshape(-1, 3)
        area = np.linalg.norm(np.cross(v[i[:, 1]] - v[i[:, 0]], v[i[:, 2]] - v[i[:, 0]]), axis=1) * 0.5
        return [DiffuseEmitter(mesh, f, radiance) for f in range(mesh.NumFaces) if area[f] != 0.0]

    Kind = 0  # sspt_emitter_kind
    Exponent = 0.0


class GlossyEmitter(DiffuseEmitter):
    """One emissive triangle whose emission is scaled by a cosine lobe about the
    shading normal, radiance * cos^exponent * (exponent + 1) / (2 pi), which keeps
    the total power of a DiffuseEmitter (SeeSharp/Shading/Emitters/GlossyEmitter.cs:7-40)."""

    Kind = 1

    def __init__(self, mesh: Mesh, face: int, radiance, exponent: float):
        super().__init__(mesh, face, radiance)
        self.Exponent = float(np.float32(exponent))

    @staticmethod
    def MakeFromMesh(mesh: Mesh, radiance, exponent: float):
        """GlossyEmitter.cs:24-34: one emitter per face, zero-area faces skipped."""
        return [GlossyEmitter(e.Mesh, e.FaceIndex, e.Radiance, exponent)
                for e in DiffuseEmitter.MakeFromMesh(mesh, radiance)]


class EnvironmentMap:
    """Lat-long RGB environment (EnvironmentMap.cs:13-17); +Y up, phi = atan2(z, x)."""

    def __init__(self, image):
        self.Image = f32(image).reshape(np.asarray(image).shape[0], np.asarray(image).shape[1], 3)


def create_look_at(position, target, up) -> np.ndarray:
    """Matrix4x4.CreateLookAt (right-handed) in fp32, row-major, row vectors."""
    p = np.asarray(position, np.float32)
    t = np.asarray(target, np.float32)
    u = np.asarray(up, np.float32)

    def norm(v):
        return (v / np.float32(np.sqrt(np.float32(v[0] * v[0] + v[1] * v[1] + v[2] * v[2])))).astype(np.float32)

    z = norm(p - t)
    x = norm(np.cross(u, z).astype(np.float32))
    y = np.cross(z, x).astype(np.float32)
    m = np.zeros((4, 4), np.float32)
    m[0:3, 0] = x
    m[0:3, 1] = y
    m[0:3, 2] = z
    m[3, 0] = -np.float32(x @ p)
    m[3, 1] = -np.float32(y @ p)
    m[3, 2] = -np.float32(z @ p)
    m[3, 3] = 1.0
    return m


class PerspectiveCamera:
    """Pinhole camera (PerspectiveCamera.cs:29-34); lens radius 0 (no depth of field)."""

    def __init__(self, worldToCamera, verticalFieldOfView: float):
        self.WorldToCamera = np.asarray(worldToCamera, np.float32).reshape(4, 4)
        if abs(np.linalg.det(self.WorldToCamera.astype(np.float64))) == 0.0:
            raise ValueError("World to camera transform must be invertible.")
        self.VerticalFieldOfView = float(verticalFieldOfView)
        self.Width = 0
        self.Height = 0

    def UpdateResolution(self, width: int, height: int):
        self.Width, self.Height = int(width), int(height)


class FrameBuffer:
    """Running-mean RGB image (SeeSharp/Images/FrameBuffer.cs). ``Image`` is H x W x 3 fp32.

    Beside the image it holds the named AOV layers, the outlier cache (per
    pixel the NumOutliersToTrack largest samples and their iterations,
    FrameBuffer.cs:29,186-189), the per-iteration error log against
    ``ReferenceImage`` (``Errors``, FrameBuffer.cs:267-291,377-383) and the
    metadata ``WriteToFile`` stores beside the image (FrameBuffer.cs:313-367)."""

    IgnoreNanAndInf = 16
    Recommended = IgnoreNanAndInf

    def __init__(self, width: int, height: int, filename: str = "", flags: int = Recommended):
        self.Width, self.Height = int(width), int(height)
        self.Filename = filename
        self.Flags = flags
        self.Image = np.zeros((self.Height, self.Width, 3), np.float32)
        self.CurIteration = 0
        self.NaNWarnings = 0
        self.MetaData = {}
        # AOV layers by name (FrameBuffer.AddLayer); the GPU path tracer fills
        # "albedo" and "normal" when EnableDenoiser is set (DenoiseBuffers).
        self.Layers = {}
        self.NumOutliersToTrack = 4
        # (weights, iterations), each H x W x NumOutliersToTrack, descending weight;
        # weight 0 / iteration -1 where a pixel has fewer samples
        self.Outliers = None
        self.ReferenceImage = None
        self.Errors: list[dict] = []

    def GetPixelOutlier(self, col: int, row: int):
        """OutlierCache.GetPixelOutlier: [(iteration, weight)] of the pixel, largest first."""
        if self.Outliers is None:
            return []
        w, it = self.Outliers
        return [(int(i), float(x)) for x, i in zip(w[row, col], it[row, col]) if i >= 0]

    def WriteToFile(self, fname: str | None = None):
        """FrameBuffer.WriteToFile (FrameBuffer.cs:313-367) for OpenEXR: the image,
        every layer and one "outlier-i" layer per tracked sample (its iteration)
        in one file, and the metadata as JSON beside it (<basename>.json)."""
        import json
        import os

        from . import imageio

        fname = fname or self.Filename
        if not fname:
            raise ValueError("no file name")
        base, ext = os.path.splitext(fname)
        if ext.lower() != ".exr":
            raise NotImplementedError(f"{fname}: only OpenEXR output is supported")
        layers = [(name, img) for name, img in self.Layers.items()] + [(None, self.Image)]
        if self.Outliers is not None:
            its = self.Outliers[1]
            for i in range(its.shape[2]):
                layers.append((f"outlier-{i}", np.maximum(its[..., i], 0).astype(np.float32)))
        imageio.write_exr_layers(fname, layers)
        meta = dict(self.MetaData)
        if self.Errors:
            meta["ErrorMetrics"] = self.Errors
        meta["NaNWarnings"] = self.NaNWarnings
        with open(base + ".json", "w") as f:
            json.dump(meta, f, indent=2)


class Scene:
    """SeeSharp/Scene.cs: meshes, emitters, background, camera and framebuffer."""

    def __init__(self):
        self.Meshes: list[Mesh] = []
        self.Emitters: list[DiffuseEmitter] = []
        self.Background: EnvironmentMap | None = None
        self.Camera: PerspectiveCamera | None = None
        self.FrameBuffer: FrameBuffer | None = None
        self.Name = ""
        self.ValidationErrorMessages: list[str] = []
        self._keep = None

    @property
    def IsValid(self) -> bool:
        """Scene.IsValid (SeeSharp/Scene.cs:170-197)."""
        msgs = []
        if self.FrameBuffer is None:
            msgs.append("Framebuffer not set.")
        if self.Camera is None:
            msgs.append("Camera not set.")
        if not self.Meshes:
            msgs.append("No meshes in the scene.")
        for i, m in enumerate(self.Meshes):
            if m.Material is None:
                msgs.append(f"Mesh[{i}] does not have a material.")
        if not self.Emitters and self.Background is None:
            msgs.append("No emitters and no background in the scene.")
        self.ValidationErrorMessages = msgs
        return not msgs

    @property
    def NumTriangles(self) -> int:
        return sum(m.NumFaces for m in self.Meshes)

    def to_desc(self) -> abi.SceneDesc:
        """Flattens the scene into an sspt_scene_desc (arrays stay alive with the scene)."""
        if not self.IsValid:
            raise ValueError("Cannot finalize an invalid scene: " + "; ".join(self.ValidationErrorMessages))
        keep = []
        materials, mat_index = [], {}
        textures, tex_index = [], {}
        tex_objects = []

        def tex_id(tex, channels):
            if tex.Image is None:
                return -1
            key = id(tex)
            if key not in tex_index:
                img = np.ascontiguousarray(tex.Image, dtype=np.float32)
                keep.append(img)
                td = abi.TextureDesc(img.shape[1], img.shape[0], channels, tex.Border, abi.fptr(img))
                tex_index[key] = len(textures)
                textures.append(td)
                tex_objects.append(tex)
            return tex_index[key]

        def mat_id(mat):
            key = id(mat)
            if key in mat_index:
                return mat_index[key]
            p = mat.MaterialParameters
            md = abi.MaterialDesc()
            bc = _tex_rgb(p.BaseColor)
            md.base_color[:] = bc.constColor
            md.base_color_texture = tex_id(bc, 3)
            if isinstance(mat, GenericMaterial):
                md.kind = abi.MATERIAL_GENERIC
                rt = _tex_mono(p.Roughness)
                md.roughness = rt.constColor
                md.roughness_texture = tex_id(rt, 1)
                md.metallic = p.Metallic
                md.specular_tint_strength = p.SpecularTintStrength
                md.anisotropic = p.Anisotropic
                md.specular_transmittance = p.SpecularTransmittance
                md.index_of_refraction = p.IndexOfRefraction
            elif isinstance(mat, DiffuseMaterial):
                md.kind = abi.MATERIAL_DIFFUSE
                md.roughness_texture = -1
                md.roughness = 1.0
                md.transmitter = 1 if p.Transmitter else 0
            else:
                raise TypeError(f"unsupported material {type(mat).__name__}")
            mat_index[key] = len(materials)
            materials.append(md)
            return mat_index[key]

        meshes = (abi.MeshDesc * len(self.Meshes))()
        mesh_index = {}
        for k, m in enumerate(self.Meshes):
            mesh_index[id(m)] = k
            keep.extend([m.Vertices, m.Indices, m.ShadingNormals, m.TextureCoordinates])
            meshes[k] = abi.MeshDesc(
                m.Vertices.shape[0], abi.fptr(m.Vertices), abi.fptr(m.ShadingNormals),
                abi.fptr(m.TextureCoordinates), m.NumFaces, abi.iptr(m.Indices), mat_id(m.Material))
        emitters = (abi.EmitterDesc * max(1, len(self.Emitters)))()
        for k, e in enumerate(self.Emitters):
            emitters[k].mesh = mesh_index[id(e.Mesh)]
            emitters[k].face = e.FaceIndex
            emitters[k].radiance[:] = e.Radiance
            emitters[k].kind = e.Kind
            emitters[k].exponent = e.Exponent
        mats = (abi.MaterialDesc * max(1, len(materials)))(*materials)
        texs = (abi.TextureDesc * max(1, len(textures)))(*textures)
        d = abi.SceneDesc()
        d.num_meshes = len(self.Meshes)
        d.meshes = meshes
        d.num_materials = len(materials)
        d.materials = mats
        d.num_textures = len(textures)
        d.textures = texs
        # texture objects in device order: the index the gradient / texture-write calls take
        self.DeviceTextures = tex_objects
        d.num_emitters = len(self.Emitters)
        d.emitters = emitters
        if self.Background is not None:
            img = np.ascontiguousarray(self.Background.Image, dtype=np.float32)
            keep.append(img)
            env = abi.EnvmapDesc(img.shape[1], img.shape[0], abi.fptr(img))
            keep.append(env)
            d.background = C.pointer(env)
        d.camera.world_to_camera[:] = [float(x) for x in self.Camera.WorldToCamera.reshape(16)]
        d.camera.vertical_fov_degrees = self.Camera.VerticalFieldOfView
        d.width = self.FrameBuffer.Width
        d.height = self.FrameBuffer.Height
        keep.extend([meshes, emitters, mats, texs])
        self._keep = keep
        return d
