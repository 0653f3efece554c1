"""Scene ingestion on the CPU: the .json scene format (SeeSharp/IO/JsonScene.cs)
with the reference's own scene files (tests/golden/scenes, copied from
Data/Scenes), .obj / .ply meshes, and the PNG / OpenEXR readers."""
import math
import os
import struct
import zlib

import numpy as np
import pytest

from seesharp_amd import DiffuseEmitter, DiffuseMaterial, GenericMaterial, GlossyEmitter, TextureRgb, imageio, sceneio

SCENES = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "scenes")


def rot_x(a):
    c, s = math.cos(a), math.sin(a)
    return np.array([[1, 0, 0, 0], [0, c, s, 0], [0, -s, c, 0], [0, 0, 0, 1]])


def rot_y(a):
    c, s = math.cos(a), math.sin(a)
    return np.array([[c, 0, -s, 0], [0, 1, 0, 0], [s, 0, c, 0], [0, 0, 0, 1]])


def rot_z(a):
    c, s = math.cos(a), math.sin(a)
    return np.array([[c, s, 0, 0], [-s, c, 0, 0], [0, 0, 1, 0], [0, 0, 0, 1]])


def test_yaw_pitch_roll_is_roll_then_pitch_then_yaw():
    """Row vectors: roll about z, then pitch about x, then yaw about y, each the
    closed-form axis rotation of System.Numerics (CreateRotationX/Y/Z)."""
    rng = np.random.default_rng(0)
    for _ in range(20):
        y, p, r = rng.uniform(-math.pi, math.pi, 3)
        m = sceneio.mat_yaw_pitch_roll(y, p, r)
        ref = rot_z(r) @ rot_x(p) @ rot_y(y)
        assert np.abs(m - ref).max() < 2e-6
    # 90 degree yaw turns +x into -z
    assert np.allclose(np.array([1, 0, 0, 0]) @ sceneio.mat_yaw_pitch_roll(math.pi / 2, 0, 0), [0, 0, -1, 0],
                       atol=1e-6)


def test_transform_is_scale_rotation_translation():
    t = {"scale": [2, 3, 4], "rotation": [0, 90, 0], "position": [1, 2, 3]}
    m = sceneio.read_transform(t)
    p = np.array([1.0, 0.0, 0.0, 1.0]) @ m  # (2,0,0) -> yaw 90 -> (0,0,-2) -> + (1,2,3)
    assert np.allclose(p, [1, 2, 1, 1], atol=1e-5)
    assert np.allclose(sceneio.read_transform({"position": [1, 2, 3], "matrix": list(range(16))}),
                       np.arange(16, dtype=np.float32).reshape(4, 4))


def test_cornell_box_json():
    s = sceneio.load_scene(os.path.join(SCENES, "CornellBox", "CornellBox.json"), 64, 48)
    assert s.IsValid
    assert len(s.Meshes) == 8 and s.NumTriangles == 36
    assert len(s.Emitters) == 2
    assert all(isinstance(m.Material, DiffuseMaterial) for m in s.Meshes)
    assert s.Meshes[0].Name == "mesh0" and s.Meshes[5].Material.MaterialParameters.BaseColor == (0.63, 0.065, 0.05)
    cam_to_world = np.linalg.inv(s.Camera.WorldToCamera.astype(np.float64))
    assert np.allclose(cam_to_world[3, :3], [0.0, 1.0, 6.8], atol=1e-5)
    assert s.Camera.VerticalFieldOfView == pytest.approx(19.5)
    assert s.FrameBuffer.Width == 64 and s.FrameBuffer.Height == 48


def test_simpledi_camera_looks_down():
    s = sceneio.load_scene(os.path.join(SCENES, "simpledi.json"))
    cam_to_world = np.linalg.inv(s.Camera.WorldToCamera.astype(np.float64))
    view = np.array([0.0, 0.0, -1.0, 0.0]) @ cam_to_world  # cameras look along -z
    assert np.allclose(view[:3], [0, -1, 0], atol=1e-6)
    assert isinstance(s.Meshes[1].Material, GenericMaterial)
    p = s.Meshes[1].Material.MaterialParameters
    assert p.IndexOfRefraction == pytest.approx(1.01) and p.SpecularTintStrength == 0.0 and p.Roughness == 0.5


def test_background_and_obj_scenes():
    s = sceneio.load_scene(os.path.join(SCENES, "simplebackground.json"))
    assert s.IsValid and s.Background is not None and s.Background.Image.shape == (900, 1600, 3)
    t = sceneio.load_scene(os.path.join(SCENES, "TextureTest", "TextureTest.json"))
    assert t.IsValid and len(t.Meshes) == 3 and len(t.Emitters) == 2
    # three .obj objects; two use the .json's textured "Material.002", one the emissive "Material.001"
    assert [m.Material.Name for m in t.Meshes] == ["Material.002", "Material.001", "Material.002"]
    textured = [m for m in t.Meshes if isinstance(m.Material.MaterialParameters.BaseColor, TextureRgb)]
    assert len(textured) == 2 and textured[0].Material is textured[1].Material
    assert all(m.TextureCoordinates is not None for m in textured)
    assert textured[0].Material.MaterialParameters.BaseColor.Image.shape == (1024, 1024, 3)
    # the furnace box alone has neither emitters nor a background (Scene.IsValid)
    f = sceneio.load_scene(os.path.join(SCENES, "furnacebox.json"))
    assert not f.IsValid


def test_obj_triangulation_and_uv_flip(tmp_path):
    (tmp_path / "q.obj").write_text(
        "v 0 0 0\nv 1 0 0\nv 1 1 0\nv 0 1 0\nvt 0 0\nvt 1 0\nvt 1 1\nvt 0 1\n"
        "usemtl A\nf 1/1 2/2 3/3 4/4\no second\nusemtl B\nf -4/-4 -3/-3 -2/-2\n")
    named = {"A": GenericMaterial(GenericMaterial.Parameters()), "B": DiffuseMaterial(DiffuseMaterial.Parameters())}
    meshes, emitters = sceneio.load_obj(str(tmp_path / "q.obj"), named, {"B": ((1.0, 2.0, 3.0), False, 50.0)})
    assert len(meshes) == 2 and len(emitters) == 1
    a, b = meshes
    assert a.Material is named["A"] and b.Material is named["B"]
    assert a.Indices.tolist() == [0, 1, 2, 0, 2, 3]
    assert np.allclose(a.TextureCoordinates[:, 1], [1, 1, 0, 0])  # v flipped
    assert np.allclose(b.Vertices, [[0, 0, 0], [1, 0, 0], [1, 1, 0]])


def test_ply_ascii_and_binary(tmp_path):
    pts = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0]], np.float32)
    head = ("ply\nformat {}\nelement vertex 4\nproperty float x\nproperty float y\nproperty float z\n"
            "element face 1\nproperty list uchar int vertex_indices\nend_header\n")
    (tmp_path / "a.ply").write_text(head.format("ascii 1.0") + "".join(f"{x} {y} {z}\n" for x, y, z in pts) +
                                    "4 0 1 2 3\n")
    body = pts.astype("<f4").tobytes() + struct.pack("<B4i", 4, 0, 1, 2, 3)
    (tmp_path / "b.ply").write_bytes(head.format("binary_little_endian 1.0").encode() + body)
    for f in ("a.ply", "b.ply"):
        p, ind, nrm, uv = sceneio.read_ply(str(tmp_path / f))
        assert np.allclose(p, pts) and ind.tolist() == [0, 1, 2, 0, 2, 3] and nrm is None and uv is None


def _png_filter(rows: np.ndarray, ftype: int, bpp: int) -> np.ndarray:
    """Encoder-side filters, written from the PNG specification's equations."""
    out = np.zeros_like(rows, dtype=np.int64)
    h, n = rows.shape
    r = rows.astype(np.int64)
    for y in range(h):
        for i in range(n):
            a = r[y, i - bpp] if i >= bpp else 0
            b = r[y - 1, i] if y > 0 else 0
            c = r[y - 1, i - bpp] if (y > 0 and i >= bpp) else 0
            pred = [0, a, b, (a + b) // 2][ftype] if ftype < 4 else None
            if ftype == 4:
                p = a + b - c
                pa, pb, pc = abs(p - a), abs(p - b), abs(p - c)
                pred = a if (pa <= pb and pa <= pc) else (b if pb <= pc else c)
            out[y, i] = (r[y, i] - pred) & 0xFF
    return out.astype(np.uint8)


@pytest.mark.parametrize("ftype", [0, 1, 2, 3, 4])
def test_png_unfilter_inverts_each_filter(ftype):
    rng = np.random.default_rng(ftype)
    rows = rng.integers(0, 256, (7, 5 * 3), dtype=np.uint8)
    filt = _png_filter(rows, ftype, 3)
    raw = b"".join(bytes([ftype]) + filt[y].tobytes() for y in range(7))
    assert np.array_equal(imageio.png_unfilter(raw, 5, 7, 3), rows)


def test_png_file_and_srgb(tmp_path):
    rng = np.random.default_rng(3)
    img = rng.integers(0, 256, (6, 4, 3), dtype=np.uint8)
    raw = b"".join(b"\x00" + img[y].tobytes() for y in range(6))

    def chunk(kind, body):
        return struct.pack(">I", len(body)) + kind + body + struct.pack(">I", zlib.crc32(kind + body) & 0xFFFFFFFF)

    data = (b"\x89PNG\r\n\x1a\n" + chunk(b"IHDR", struct.pack(">IIBBBBB", 4, 6, 8, 2, 0, 0, 0)) +
            chunk(b"IDAT", zlib.compress(raw)) + chunk(b"IEND", b""))
    (tmp_path / "t.png").write_bytes(data)
    lin = imageio.read_png(str(tmp_path / "t.png"))
    raw_v = imageio.read_png(str(tmp_path / "t.png"), linear=False)
    assert np.allclose(raw_v, img / 255.0)
    v = img / 255.0
    assert np.allclose(lin, np.where(v <= 0.04045, v / 12.92, ((v + 0.055) / 1.055) ** 2.4), atol=1e-6)


@pytest.mark.parametrize("compression", ["none", "zip"])
def test_exr_round_trip(tmp_path, compression):
    rng = np.random.default_rng(5)
    img = (rng.random((35, 21, 3)) * 100).astype(np.float32)
    f = str(tmp_path / "t.exr")
    imageio.write_exr(f, img, compression)
    assert np.array_equal(imageio.read_exr(f), img)


def test_reference_exr_loads():
    sky = imageio.read_exr(os.path.join(SCENES, "sunsky.exr"))
    assert sky.shape == (900, 1600, 3) and np.all(np.isfinite(sky))
    assert sky.min() >= 0 and sky.max() > 100  # the sun


def test_glossy_emissive_material(tmp_path):
    """emissionIsGlossy / emissionExponent (JsonScene.cs:93-104, exponent default
    50) make GlossyEmitters; plain emission makes DiffuseEmitters."""
    scene = {
        "transforms": [{"name": "cam", "position": [0, 0, 3]}],
        "cameras": [{"name": "c", "type": "perspective", "fov": 40, "transform": "cam"}],
        "materials": [
            {"name": "g", "baseColor": {"type": "rgb", "value": [0, 0, 0]},
             "emission": {"type": "rgb", "value": [5, 5, 5]}, "emissionIsGlossy": True, "emissionExponent": 12},
            {"name": "h", "baseColor": {"type": "rgb", "value": [0, 0, 0]},
             "emission": {"type": "rgb", "value": [1, 2, 3]}, "emissionIsGlossy": True},
            {"name": "d", "baseColor": {"type": "rgb", "value": [0, 0, 0]},
             "emission": {"type": "rgb", "value": [7, 7, 7]}},
        ],
        "objects": [
            {"name": n, "material": n, "type": "trimesh", "indices": [0, 1, 2],
             "vertices": [0, 0, z, 1, 0, z, 0, 1, z]} for n, z in (("g", 0), ("h", -1), ("d", -2))
        ],
    }
    import json

    (tmp_path / "s.json").write_text(json.dumps(scene))
    s = sceneio.load_scene(str(tmp_path / "s.json"), 8, 8)
    g, h, d = s.Emitters
    assert isinstance(g, GlossyEmitter) and g.Exponent == 12.0 and g.Radiance == (5.0, 5.0, 5.0)
    assert isinstance(h, GlossyEmitter) and h.Exponent == 50.0
    assert type(d) is DiffuseEmitter and d.Radiance == (7.0, 7.0, 7.0)
