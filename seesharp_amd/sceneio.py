"""Scene ingestion: SeeSharp's .json scene format with its "trimesh", "obj" and
"ply" objects, turned into the host scene model (scene.py) that GpuScene
uploads. Mirrors JsonScene.LoadFromFile (SeeSharp/IO/JsonScene.cs:189-216),
TriMeshLoader (SeeSharp/IO/TriMeshLoader.cs:8-64), ObjMesh / ObjConverter
(SeeSharp/IO/ObjMesh.cs, ObjConverter.cs) and PlyFile (SeeSharp/IO/PlyFile.cs)
in behaviour:

- transforms: scale, then rotation (yaw = y, pitch = x, roll = z degrees, as
  Matrix4x4.CreateFromYawPitchRoll), then translation, composed for row
  vectors (v * S * R * T); "matrix" (9, 12 or 16 entries) replaces them
  (JsonScene.cs:143-182, JsonUtils.GetMatrix);
- cameras: world-to-camera = inverse of the named transform, vertical fov in
  degrees; the last camera listed is the scene's (JsonScene.cs:122-134);
- materials: "diffuse" -> DiffuseMaterial (baseColor, "thin" -> Transmitter),
  anything else -> GenericMaterial with the .json defaults roughness 0.5,
  IOR 1.01, specularTint 0, anisotropic 0, metallic 0, specularTransmittance 0
  (JsonScene.cs:70-88); a non-black "emission" makes every mesh using the
  material an area light; an object-level "emission" does the same for one
  trimesh object (TriMeshLoader.cs:53-60);
- .obj: faces fan-triangulated, one mesh per (object, group, material) in
  order of appearance, texture v flipped (1 - v), .json materials replace
  equally named .obj materials (ObjConverter.cs:52-205);
- background {"type": "image"} -> EnvironmentMap of the image.

Images are read by seesharp_amd.imageio. Emissive materials make a
DiffuseEmitter or, with "emissionIsGlossy", a GlossyEmitter per triangle.
"""
from __future__ import annotations

import json
import math
import os
import struct

import numpy as np

from . import imageio
from .scene import (
    DiffuseEmitter,
    DiffuseMaterial,
    EnvironmentMap,
    FrameBuffer,
    GenericMaterial,
    GlossyEmitter,
    Mesh,
    PerspectiveCamera,
    Scene,
    TextureMono,
    TextureRgb,
)

# ------------------------------------------------------------ Matrix4x4 ------
# System.Numerics.Matrix4x4 conventions: row vectors (p' = p * M), translation
# in the fourth row, products in float32.


def mat_scale(s) -> np.ndarray:
    m = np.eye(4, dtype=np.float32)
    m[0, 0], m[1, 1], m[2, 2] = np.float32(s[0]), np.float32(s[1]), np.float32(s[2])
    return m


def mat_translation(p) -> np.ndarray:
    m = np.eye(4, dtype=np.float32)
    m[3, :3] = np.asarray(p, np.float32)
    return m


def mat_yaw_pitch_roll(yaw: float, pitch: float, roll: float) -> np.ndarray:
    """Rotation by roll about z, then pitch about x, then yaw about y (row
    vectors), built from the unit quaternion of the three half angles."""
    f = np.float32
    sr, cr = f(math.sin(roll * 0.5)), f(math.cos(roll * 0.5))
    sp, cp = f(math.sin(pitch * 0.5)), f(math.cos(pitch * 0.5))
    sy, cy = f(math.sin(yaw * 0.5)), f(math.cos(yaw * 0.5))
    x = cy * sp * cr + sy * cp * sr
    y = sy * cp * cr - cy * sp * sr
    z = cy * cp * sr - sy * sp * cr
    w = cy * cp * cr + sy * sp * sr
    xx, yy, zz = x * x, y * y, z * z
    xy, wz, xz, wy, yz, wx = x * y, w * z, x * z, w * y, y * z, w * x
    one, two = f(1), f(2)
    m = np.eye(4, dtype=np.float32)
    m[0, :3] = [one - two * (yy + zz), two * (xy + wz), two * (xz - wy)]
    m[1, :3] = [two * (xy - wz), one - two * (zz + xx), two * (yz + wx)]
    m[2, :3] = [two * (xz + wy), two * (yz - wx), one - two * (yy + xx)]
    return m


def mat_from_entries(v) -> np.ndarray:
    """JsonUtils.GetMatrix: 9 (3x3), 12 (3x4) or 16 (4x4) row-major entries."""
    v = [float(x) for x in v]
    if len(v) not in (9, 12, 16):
        raise ValueError(f"a matrix has 9, 12 or 16 entries, not {len(v)}")
    m = np.eye(4, dtype=np.float32)
    m[0, :3] = v[0:3]
    m[1, :3] = v[4:7] if len(v) > 9 else v[3:6]
    m[2, :3] = v[8:11] if len(v) > 9 else v[6:9]
    if len(v) >= 12:
        m[0, 3], m[1, 3], m[2, 3] = v[3], v[7], v[11]
    if len(v) == 16:
        m[3, :] = v[12:16]
    return m


def read_transform(t: dict) -> np.ndarray:
    m = np.eye(4, dtype=np.float32)
    if "scale" in t:
        m = (m @ mat_scale(t["scale"])).astype(np.float32)
    if "rotation" in t:
        r = [math.radians(a) for a in t["rotation"]]
        m = (m @ mat_yaw_pitch_roll(r[1], r[0], r[2])).astype(np.float32)
    if "position" in t:
        m = (m @ mat_translation(t["position"])).astype(np.float32)
    if "matrix" in t:
        m = mat_from_entries(t["matrix"])
    return m


def invert(m: np.ndarray) -> np.ndarray:
    return np.linalg.inv(m.astype(np.float64)).astype(np.float32)


# ------------------------------------------------------------- materials -----
def _color_or_texture(spec: dict, base_dir: str):
    kind = spec.get("type")
    if kind == "rgb":
        return tuple(float(x) for x in spec["value"][:3])
    if kind == "image":
        return TextureRgb(imageio.read_image(os.path.join(base_dir, spec["filename"])))
    raise ValueError(f"invalid texture specification {spec}")


def _float_or_texture(m: dict, name: str, base_dir: str, default: float):
    if name not in m:
        return default
    v = m[name]
    if isinstance(v, (int, float)):
        return float(v)
    img = imageio.read_image(os.path.join(base_dir, v))
    return TextureMono(img.mean(axis=-1).astype(np.float32))


def read_materials(root: dict, base_dir: str):
    named, emissive = {}, {}
    for m in root.get("materials", []):
        name = m["name"]
        if m.get("type", "generic") == "diffuse":
            mat = DiffuseMaterial(DiffuseMaterial.Parameters(
                BaseColor=_color_or_texture(m["baseColor"], base_dir), Transmitter=bool(m.get("thin", False))))
        else:
            mat = GenericMaterial(GenericMaterial.Parameters(
                BaseColor=_color_or_texture(m["baseColor"], base_dir),
                Roughness=_float_or_texture(m, "roughness", base_dir, 0.5),
                Anisotropic=float(m.get("anisotropic", 0.0)),
                IndexOfRefraction=float(m.get("IOR", 1.01)),
                Metallic=float(m.get("metallic", 0.0)),
                SpecularTintStrength=float(m.get("specularTint", 0.0)),
                SpecularTransmittance=float(m.get("specularTransmittance", 0.0))))
        mat.Name = name
        named[name] = mat
        if "emission" in m:
            e = tuple(float(x) for x in m["emission"]["value"][:3])
            if any(c != 0.0 for c in e):
                # EmissionParameters (JsonScene.cs:93-104): glossy lobe exponent defaults to 50
                emissive[name] = (e, bool(m.get("emissionIsGlossy", False)), float(m.get("emissionExponent", 50.0)))
    return named, emissive


# ---------------------------------------------------------------- meshes -----
def _f32(v, n):
    a = np.asarray(v, np.float32)
    return a.reshape(-1, n)


def _material_emitters(mesh: Mesh, params):
    """Emitters of a mesh whose material is emissive (TriMeshLoader.cs:53-56,
    PlyLoader.cs:31-34, ObjConverter.cs:200-203): glossy or diffuse per the
    material's emission parameters."""
    radiance, glossy, exponent = params
    if glossy:
        return GlossyEmitter.MakeFromMesh(mesh, radiance, exponent)
    return DiffuseEmitter.MakeFromMesh(mesh, radiance)


def load_trimesh(obj: dict, named, emissive):
    """TriMeshLoader: inline vertices / indices / normals / uv arrays."""
    name = obj["material"]
    mesh = Mesh(_f32(obj["vertices"], 3), np.asarray(obj["indices"], np.int32),
                _f32(obj["normals"], 3) if "normals" in obj else None,
                _f32(obj["uv"], 2) if "uv" in obj else None)
    mesh.Material = named[name]
    emitters = []
    if name in emissive:
        emitters = _material_emitters(mesh, emissive[name])
    elif "emission" in obj:
        emitters = DiffuseEmitter.MakeFromMesh(mesh, tuple(float(x) for x in obj["emission"]["value"][:3]))
    return [mesh], emitters


def _read_mtl(path: str):
    """Minimal .mtl reader: Kd, Ks, Ke, Ns, Ni, illum, map_Kd per material."""
    mats, cur = {}, None
    if not os.path.exists(path):
        return mats
    for line in open(path, errors="replace"):
        tok = line.strip().split()
        if not tok or tok[0].startswith("#"):
            continue
        if tok[0] == "newmtl":
            cur = {"Kd": (0.8, 0.8, 0.8), "Ks": (0.0, 0.0, 0.0), "Ke": (0.0, 0.0, 0.0), "Ns": 0.0, "Ni": 1.0,
                   "illum": 2, "map_Kd": None}
            mats[" ".join(tok[1:])] = cur
        elif cur is not None:
            if tok[0] in ("Kd", "Ks", "Ke"):
                cur[tok[0]] = tuple(float(x) for x in tok[1:4])
            elif tok[0] in ("Ns", "Ni"):
                cur[tok[0]] = float(tok[1])
            elif tok[0] == "illum":
                cur["illum"] = int(tok[1])
            elif tok[0] == "map_Kd":
                cur["map_Kd"] = tok[-1]
    return mats


def _obj_material(spec: dict, base_dir: str):
    """ObjConverter's mapping of illumination models to GenericMaterial."""
    if spec["illum"] == 5:  # perfect mirror
        p = GenericMaterial.Parameters(BaseColor=spec["Ks"], SpecularTintStrength=1.0, Metallic=1.0, Roughness=0.0)
    elif spec["illum"] == 7:  # perfect glass
        p = GenericMaterial.Parameters(BaseColor=spec["Ks"], Metallic=0.0, Roughness=0.0,
                                       IndexOfRefraction=spec["Ni"], SpecularTransmittance=1.0,
                                       SpecularTintStrength=1.0)
    else:  # treated as Phong: diffuse colour or texture, roughness from the exponent
        base = spec["Kd"] if not spec["map_Kd"] else TextureRgb(
            imageio.read_image(os.path.join(base_dir, spec["map_Kd"])))
        p = GenericMaterial.Parameters(BaseColor=base, Roughness=1.0 if spec["Ns"] == 0 else 1.0 / spec["Ns"],
                                       Metallic=0.5)
    return GenericMaterial(p)


def load_obj(path: str, named, emissive):
    base_dir = os.path.dirname(path)
    verts, norms, uvs = [np.zeros(3, np.float32)], [np.zeros(3, np.float32)], [np.zeros(2, np.float32)]
    objects = [[[]]]  # objects -> groups -> faces (material index, [(v, t, n), ...])
    mat_names = [""]
    mtl_files = []
    cur_mtl = 0
    for line in open(path, errors="replace"):
        line = line.strip()
        if not line or line[0] == "#":
            continue
        tok = line.split()
        head = tok[0]
        if head == "v":
            verts.append(np.asarray([float(x) for x in tok[1:4]], np.float32))
        elif head == "vn":
            norms.append(np.asarray([float(x) for x in tok[1:4]], np.float32))
        elif head == "vt":
            uvs.append(np.asarray([float(x) for x in tok[1:3]], np.float32))
        elif head == "f":
            idx = []
            for t in tok[1:]:
                parts = (t.split("/") + ["", ""])[:3]
                v, vt, vn = (int(p) if p else 0 for p in parts)
                v = len(verts) + v if v < 0 else v
                vt = len(uvs) + vt if vt < 0 else vt
                vn = len(norms) + vn if vn < 0 else vn
                idx.append((v, vt, vn))
            if len(idx) >= 3 and all(v > 0 and vt >= 0 and vn >= 0 for v, vt, vn in idx):
                objects[-1][-1].append((cur_mtl, idx))
        elif head == "g":
            objects[-1].append([])
        elif head == "o":
            objects.append([[]])
        elif head == "usemtl":
            name = line[6:].strip()
            if name not in mat_names:
                mat_names.append(name)
            cur_mtl = mat_names.index(name)
        elif head == "mtllib":
            mtl_files.append(line[6:].strip())
    lib = {}
    for f in mtl_files:
        lib.update(_read_mtl(os.path.join(base_dir, f)))
    white = GenericMaterial(GenericMaterial.Parameters(BaseColor=(1.0, 1.0, 1.0)))
    materials, obj_emission = {}, {}
    for name in mat_names:
        if name in named:
            materials[name] = named[name]
        elif name in lib:
            materials[name] = _obj_material(lib[name], base_dir)
            if any(c != 0.0 for c in lib[name]["Ke"]):
                obj_emission[name] = lib[name]["Ke"]
        else:
            materials[name] = white
    meshes, emitters = [], []
    for groups in objects:
        faces = [f for g in groups for f in g]
        if not faces:
            continue
        has_n = any(vn != 0 for _, idx in faces for _, _, vn in idx)
        has_t = any(vt != 0 for _, idx in faces for _, vt, _ in idx)
        for g in groups:
            by_mtl = {}
            for m, idx in g:
                tris = by_mtl.setdefault(m, [])
                for i in range(1, len(idx) - 1):
                    tris.append((idx[0], idx[i], idx[i + 1]))
            for m, tris in by_mtl.items():
                remap, lv, ln, lt, ind = {}, [], [], [], []
                for tri in tris:
                    for key in tri:
                        if key not in remap:
                            remap[key] = len(lv)
                            lv.append(verts[key[0]])
                            if has_n:
                                ln.append(norms[key[2]])
                            if has_t:
                                u = uvs[key[1]].copy()
                                u[1] = np.float32(1.0) - u[1]
                                lt.append(u)
                        ind.append(remap[key])
                mesh = Mesh(np.asarray(lv, np.float32), np.asarray(ind, np.int32),
                            np.asarray(ln, np.float32) if has_n else None,
                            np.asarray(lt, np.float32) if has_t else None)
                name = mat_names[m]
                mesh.Material = materials[name]
                meshes.append(mesh)
                if name in emissive:
                    emitters += _material_emitters(mesh, emissive[name])
                elif name in obj_emission:
                    emitters += DiffuseEmitter.MakeFromMesh(mesh, obj_emission[name])
    return meshes, emitters


_PLY_TYPES = {"char": "i1", "int8": "i1", "uchar": "u1", "uint8": "u1", "short": "i2", "int16": "i2",
              "ushort": "u2", "uint16": "u2", "int": "i4", "int32": "i4", "uint": "u4", "uint32": "u4",
              "float": "f4", "float32": "f4", "double": "f8", "float64": "f8"}


def read_ply(path: str):
    """Triangle / polygon .ply (ascii or binary): positions, optional normals
    (nx, ny, nz) and uvs (u, v or s, t), faces fan-triangulated."""
    data = open(path, "rb").read()
    end = data.index(b"end_header") + len(b"end_header")
    end = data.index(b"\n", end) + 1
    header = data[:end].decode(errors="replace").splitlines()
    if header[0].strip() != "ply":
        raise ValueError(f"{path}: not a PLY file")
    fmt, elements = None, []
    for line in header[1:]:
        tok = line.split()
        if not tok:
            continue
        if tok[0] == "format":
            fmt = tok[1]
        elif tok[0] == "element":
            elements.append([tok[1], int(tok[2]), []])
        elif tok[0] == "property":
            elements[-1][2].append(tok[1:])
    body = data[end:]
    order = {"binary_little_endian": "<", "binary_big_endian": ">"}.get(fmt)
    values = {}
    pos = 0
    words = body.split() if order is None else None
    wpos = 0
    for name, count, props in elements:
        rows = []
        for _ in range(count):
            row = {}
            for p in props:
                if p[0] == "list":
                    ct, it, pname = _PLY_TYPES[p[1]], _PLY_TYPES[p[2]], p[3]
                    if order is None:
                        n = int(words[wpos]); wpos += 1
                        row[pname] = [float(w) for w in words[wpos:wpos + n]]; wpos += n
                    else:
                        n = struct.unpack_from(order + np.dtype(ct).char, body, pos)[0]
                        pos += np.dtype(ct).itemsize
                        row[pname] = list(np.frombuffer(body, np.dtype(it).newbyteorder(order), n, pos))
                        pos += n * np.dtype(it).itemsize
                else:
                    t, pname = _PLY_TYPES[p[0]], p[1]
                    if order is None:
                        row[pname] = float(words[wpos]); wpos += 1
                    else:
                        row[pname] = struct.unpack_from(order + np.dtype(t).char, body, pos)[0]
                        pos += np.dtype(t).itemsize
            rows.append(row)
        values[name] = rows
    vs = values.get("vertex", [])
    pts = np.asarray([[r["x"], r["y"], r["z"]] for r in vs], np.float32)
    nrm = np.asarray([[r["nx"], r["ny"], r["nz"]] for r in vs], np.float32) if vs and "nx" in vs[0] else None
    uv = None
    if vs and ("u" in vs[0] or "s" in vs[0]):
        ku, kv = ("u", "v") if "u" in vs[0] else ("s", "t")
        uv = np.asarray([[r[ku], r[kv]] for r in vs], np.float32)
    ind = []
    for r in values.get("face", []):
        f = r.get("vertex_indices", r.get("vertex_index"))
        for i in range(1, len(f) - 1):
            ind += [int(f[0]), int(f[i]), int(f[i + 1])]
    return pts, np.asarray(ind, np.int32), nrm, uv


def load_ply(path: str, obj: dict, named, emissive):
    pts, ind, nrm, uv = read_ply(path)
    mesh = Mesh(pts, ind, nrm, uv)
    name = obj["material"]
    mesh.Material = named[name]
    emitters = []
    if name in emissive:
        emitters = _material_emitters(mesh, emissive[name])
    elif "emission" in obj:
        emitters = DiffuseEmitter.MakeFromMesh(mesh, tuple(float(x) for x in obj["emission"]["value"][:3]))
    return [mesh], emitters


# ------------------------------------------------------------------ scene ----
def load_scene(path: str, width: int = 512, height: int = 512) -> Scene:
    """JsonScene.LoadFromFile plus a width x height FrameBuffer."""
    root = json.load(open(path))
    base_dir = os.path.dirname(os.path.abspath(path))
    transforms = {t["name"]: read_transform(t) for t in root.get("transforms", [])}
    s = Scene()
    s.Name = root.get("name", os.path.basename(path))
    for c in root.get("cameras", []):
        if c.get("type", "perspective") != "perspective":
            raise NotImplementedError(f"camera type {c.get('type')}")
        s.Camera = PerspectiveCamera(invert(transforms[c["transform"]]), float(c["fov"]))
    bg = root.get("background")
    if bg and bg.get("type") == "image":
        s.Background = EnvironmentMap(imageio.read_image(os.path.join(base_dir, bg["filename"])))
    named, emissive = read_materials(root, base_dir)
    for obj in root.get("objects", []):
        kind = obj["type"]
        if kind == "trimesh":
            meshes, emitters = load_trimesh(obj, named, emissive)
        elif kind == "obj":
            meshes, emitters = load_obj(os.path.join(base_dir, obj["relativePath"]), named, emissive)
        elif kind == "ply":
            meshes, emitters = load_ply(os.path.join(base_dir, obj["relativePath"]), obj, named, emissive)
        else:
            raise NotImplementedError(f"object type {kind}")
        for k, m in enumerate(meshes):
            m.Name = obj.get("name", "") + ("" if k == 0 else f".{k - 1:03d}")
        s.Meshes.extend(meshes)
        s.Emitters.extend(emitters)
    s.RecommendedExposure = float(root.get("exposure", 0.0))
    s.FrameBuffer = FrameBuffer(width, height)
    return s
