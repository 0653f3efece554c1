"""Seeded synthetic scenes of the benchmark configurations.

Every random choice is drawn from SeeSharp's own generator
(SeeSharp/Sampling/RNG.cs), seeded as RNG(HashSeed(sceneSeed, objectIndex, k)),
so the scenes are a pure function of their seeds.

* ``box_scene``      -- 36 triangles: open unit box, two rotated boxes, a
                        0.25 x 0.25 ceiling light; 256^2, the layout of
                        Data/Scenes/CornellBox/CornellBox.json.
* ``large_scene``    -- 1 036 296 triangles: 400 icospheres (1280 tris each),
                        a 512 x 512-quad fBm ground grid, 4 quad lights and a
                        procedural 512 x 256 sky; 1024^2.
* ``textured_box_scene`` -- the box at 512^2 with albedo / roughness textures
                        on the back wall and floor.
"""
from __future__ import annotations

import math

import numpy as np

from .scene import (
    DiffuseEmitter,
    DiffuseMaterial,
    EnvironmentMap,
    FrameBuffer,
    GenericMaterial,
    Mesh,
    PerspectiveCamera,
    Scene,
    TextureMono,
    TextureRgb,
    create_look_at,
)

M32 = 0xFFFFFFFF


def pcg_hash(x: int) -> int:
    """RNG.PcgHash (SeeSharp/Sampling/RNG.cs:107-111)."""
    state = (x * 747796405 + 2891336453) & M32
    word = (((state >> ((state >> 28) + 4)) ^ state) * 277803737) & M32
    return ((word >> 22) ^ word) & M32


def fnv_hash(h: int, data: int) -> int:
    """RNG.FnvHash (RNG.cs:99-105)."""
    for shift in (0, 8, 16, 24):
        h = ((h * 16777619) & M32) ^ ((data >> shift) & 0xFF)
    return h


def hash_seed(base: int, chain: int, sample: int) -> int:
    """RNG.HashSeed (RNG.cs:113-118)."""
    h = fnv_hash(2166136261, pcg_hash(base & M32))
    h = fnv_hash(h, pcg_hash(chain & M32))
    h = fnv_hash(h, pcg_hash(sample & M32))
    return pcg_hash(h)


class Rng:
    """SeeSharp/Sampling/RNG.cs: PCG-RXS-M-XS 32 bit."""

    def __init__(self, seed: int):
        self.state = seed & M32

    @classmethod
    def seeded(cls, base: int, chain: int, sample: int) -> "Rng":
        return cls(hash_seed(base, chain, sample))

    def next(self) -> int:
        s = self.state
        word = (((s >> ((s >> 28) + 4)) ^ s) * 277803737) & M32
        self.state = (s * 747796405 + 2891336453) & M32
        return ((word >> 22) ^ word) & M32

    def next_float(self, lo: float | None = None, hi: float | None = None) -> float:
        r = np.float32(np.float32(self.next()) / np.float32(4294967296.0))
        if lo is None:
            return float(r)
        return float(np.float32(lo) * (np.float32(1) - r) + np.float32(hi) * r)


# ------------------------------------------------------------------ geometry --

def quad(p0, p1, p2, p3):
    """Two triangles (p0,p1,p2), (p0,p2,p3); face normal = normalize(cross(p1-p0, p2-p0))."""
    v = np.array([p0, p1, p2, p3], np.float32)
    return v, np.array([0, 1, 2, 0, 2, 3], np.int32)


def box_mesh(center, size, angle_deg):
    """Axis-aligned box rotated about +y, outward-facing CCW triangles (12)."""
    sx, sy, sz = (s * 0.5 for s in size)
    c, s = math.cos(math.radians(angle_deg)), math.sin(math.radians(angle_deg))
    corners = []
    for x in (-sx, sx):
        for y in (-sy, sy):
            for z in (-sz, sz):
                corners.append((x, y, z))
    faces = [  # outward normal order (verified by test_scenes)
        (0, 1, 3, 2),  # -x
        (4, 6, 7, 5),  # +x
        (0, 4, 5, 1),  # -y
        (2, 3, 7, 6),  # +y
        (0, 2, 6, 4),  # -z
        (1, 5, 7, 3),  # +z
    ]
    verts = []
    for (x, y, z) in corners:
        verts.append((center[0] + c * x + s * z, center[1] + y, center[2] - s * x + c * z))
    v = np.array(verts, np.float32)
    idx = []
    for a, b, cc, d in faces:
        idx += [a, b, cc, a, cc, d]
    return v, np.array(idx, np.int32)


def icosphere(subdivisions: int):
    """Unit icosphere: 20 * 4^subdivisions outward CCW faces, shared vertices."""
    t = (1.0 + math.sqrt(5.0)) / 2.0
    verts = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t),
             (0, -1, -t), (0, 1, -t), (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    verts = [np.array(v, np.float64) / np.linalg.norm(v) for v in verts]
    faces = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4),
             (11, 10, 2), (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8),
             (3, 8, 9), (4, 9, 5), (2, 4, 11), (6, 2, 10), (8, 6, 7), (9, 8, 1)]
    for _ in range(subdivisions):
        cache = {}

        def mid(a, b):
            key = (min(a, b), max(a, b))
            if key not in cache:
                m = verts[a] + verts[b]
                verts.append(m / np.linalg.norm(m))
                cache[key] = len(verts) - 1
            return cache[key]

        nf = []
        for a, b, c in faces:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            nf += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        faces = nf
    v = np.array(verts, np.float64)
    f = np.array(faces, np.int32)
    # make every face wind outward
    n = np.cross(v[f[:, 1]] - v[f[:, 0]], v[f[:, 2]] - v[f[:, 0]])
    flip = (n * v[f].mean(axis=1)).sum(axis=1) < 0
    f[flip] = f[flip][:, [0, 2, 1]]
    return v, f


def value_noise_fbm(x, z, seed: int, octaves: int = 5):
    """fBm of bilinear value noise; lattice values from RNG.PcgHash (deterministic)."""
    x = np.asarray(x, np.float64)
    z = np.asarray(z, np.float64)
    total = np.zeros_like(x)
    amp, freq, norm = 1.0, 0.5, 0.0

    def lattice(ix, iz, o):
        h = (ix.astype(np.int64) * 73856093) ^ (iz.astype(np.int64) * 19349663) ^ (o * 83492791 + seed)
        h = h & M32
        st = (h * 747796405 + 2891336453) & M32
        word = (((st >> ((st >> 28) + 4)) ^ st) * 277803737) & M32
        word = ((word >> 22) ^ word) & M32
        return word.astype(np.float64) / 4294967295.0 * 2.0 - 1.0

    for o in range(octaves):
        fx, fz = x * freq, z * freq
        ix, iz = np.floor(fx), np.floor(fz)
        tx, tz = fx - ix, fz - iz
        tx = tx * tx * (3 - 2 * tx)
        tz = tz * tz * (3 - 2 * tz)
        a = lattice(ix, iz, o)
        b = lattice(ix + 1, iz, o)
        c = lattice(ix, iz + 1, o)
        d = lattice(ix + 1, iz + 1, o)
        total += amp * ((a * (1 - tx) + b * tx) * (1 - tz) + (c * (1 - tx) + d * tx) * tz)
        norm += amp
        amp *= 0.5
        freq *= 2.0
    return total / norm


def ground_grid(n: int, half: float, amplitude: float, seed: int):
    """n x n quads over [-half, half]^2 with fBm heights, smooth normals, uv = grid / n."""
    g = np.linspace(-half, half, n + 1)
    X, Z = np.meshgrid(g, g, indexing="ij")  # X along i, Z along j
    H = amplitude * value_noise_fbm(X, Z, seed)
    pos = np.stack([X, H, Z], axis=-1).reshape(-1, 3)
    step = 2 * half / n
    Hp = np.pad(H, 1, mode="edge")
    dhdx = (Hp[2:, 1:-1] - Hp[:-2, 1:-1]) / (2 * step)
    dhdz = (Hp[1:-1, 2:] - Hp[1:-1, :-2]) / (2 * step)
    nrm = np.stack([-dhdx, np.ones_like(H), -dhdz], axis=-1)
    nrm /= np.linalg.norm(nrm, axis=-1, keepdims=True)
    ii, jj = np.meshgrid(np.arange(n + 1), np.arange(n + 1), indexing="ij")
    uv = np.stack([ii / n, jj / n], axis=-1).reshape(-1, 2)
    i, j = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    v00 = (i * (n + 1) + j).reshape(-1)
    v01 = v00 + 1
    v10 = v00 + (n + 1)
    v11 = v10 + 1
    # (v00, v01, v11) and (v00, v11, v10) wind to +y
    idx = np.stack([v00, v01, v11, v00, v11, v10], axis=-1).reshape(-1)
    return pos.astype(np.float32), idx.astype(np.int32), nrm.reshape(-1, 3).astype(np.float32), uv.astype(np.float32)


def light_quad(cx, y, cz, size):
    """Quad of the given size at height y facing -y."""
    h = size * 0.5
    return quad((cx - h, y, cz - h), (cx + h, y, cz - h), (cx + h, y, cz + h), (cx - h, y, cz + h))


def sky_image(width: int = 512, height: int = 256):
    """Procedural lat-long sky: vertical gradient, dim ground, a sun disc."""
    rows = (np.arange(height) + 0.5) / height * math.pi  # theta from +y
    cols = (np.arange(width) + 0.5) / width * 2 * math.pi  # phi = atan2(z, x)
    theta, phi = np.meshgrid(rows, cols, indexing="ij")
    d = np.stack([np.sin(theta) * np.cos(phi), np.cos(theta), np.sin(theta) * np.sin(phi)], -1)
    up = np.clip(d[..., 1], 0, 1)[..., None]
    zenith = np.array([0.25, 0.45, 0.9])
    horizon = np.array([0.9, 0.85, 0.8])
    ground = np.array([0.15, 0.13, 0.1])
    sky = horizon * (1 - up) + zenith * up
    img = np.where(d[..., 1:2] >= 0, sky, ground)
    sun = np.array([0.4, 0.8, 0.3])
    sun /= np.linalg.norm(sun)
    cosang = d @ sun
    img = img + (cosang > math.cos(math.radians(3.0)))[..., None] * np.array([40.0, 38.0, 32.0])
    return img.astype(np.float32)


# -------------------------------------------------------------------- scenes --

def box_scene(width: int = 256, height: int = 256, textured: bool = False, tex_res: int = 256,
              tex_seed: int = 3, uniform_textures: float | None = None, spec_tint: float = 1.0,
              diffuse_textures: bool = False) -> Scene:
    """The 36-triangle box (sceneSeed 0). `textured` gives the back wall and floor
    textured materials: GenericMaterials with albedo + roughness textures, or
    textured DiffuseMaterials with `diffuse_textures`."""
    s = Scene()
    s.Name = "box"

    def add(v, idx, mat, normals=None, uvs=None):
        m = Mesh(v, idx, normals, uvs)
        m.Material = mat
        s.Meshes.append(m)
        return m

    white = DiffuseMaterial(DiffuseMaterial.Parameters(BaseColor=(0.8, 0.8, 0.8)))
    red = DiffuseMaterial(DiffuseMaterial.Parameters(BaseColor=(0.8, 0.1, 0.1)))
    green = DiffuseMaterial(DiffuseMaterial.Parameters(BaseColor=(0.1, 0.8, 0.1)))
    grey = DiffuseMaterial(DiffuseMaterial.Parameters(BaseColor=(0.7, 0.7, 0.7)))
    black = DiffuseMaterial(DiffuseMaterial.Parameters(BaseColor=(0.0, 0.0, 0.0)))
    uv_quad = np.array([[0, 0], [1, 0], [1, 1], [0, 1]], np.float32)

    back_mat, floor_mat = white, white
    if textured:
        back_mat, floor_mat = textured_materials(tex_res, tex_seed, uniform_textures, spec_tint, diffuse_textures)

    add(*quad((0, 0, 0), (0, 0, 1), (1, 0, 1), (1, 0, 0)), floor_mat, None, uv_quad)      # floor, +y
    add(*quad((0, 1, 0), (1, 1, 0), (1, 1, 1), (0, 1, 1)), white)                           # ceiling, -y
    add(*quad((0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0)), back_mat, None, uv_quad)        # back, +z
    add(*quad((0, 0, 0), (0, 1, 0), (0, 1, 1), (0, 0, 1)), red)                             # left, +x
    add(*quad((1, 0, 0), (1, 0, 1), (1, 1, 1), (1, 1, 0)), green)                           # right, -x
    add(*box_mesh((0.68, 0.15, 0.62), (0.3, 0.3, 0.3), 18.0), grey)
    add(*box_mesh((0.32, 0.3, 0.35), (0.3, 0.6, 0.3), -18.0), grey)
    light = add(*light_quad(0.5, 1.0 - 1e-3, 0.5, 0.25), black)
    s.Emitters.extend(DiffuseEmitter.MakeFromMesh(light, (15.0, 15.0, 15.0)))
    s.Camera = PerspectiveCamera(create_look_at((0.5, 0.5, 2.4), (0.5, 0.5, 0.0), (0, 1, 0)), 40.0)
    s.FrameBuffer = FrameBuffer(width, height)
    return s


def smooth_noise_texture(res: int, channels: int, seed: int, lo: float, hi: float):
    x = np.linspace(0, 4, res, endpoint=False)
    X, Z = np.meshgrid(x, x, indexing="ij")
    out = []
    for c in range(channels):
        n = value_noise_fbm(X, Z, seed * 7 + c, octaves=3)
        out.append(lo + (hi - lo) * (0.5 + 0.5 * n))
    img = np.stack(out, -1).astype(np.float32)
    return img if channels > 1 else img[..., 0]


def textured_materials(res: int, seed: int, uniform: float | None, spec_tint: float = 1.0,
                       diffuse: bool = False):
    """Materials with res^2 RGB albedo and (GenericMaterial) mono roughness textures
    for the back wall and the floor."""
    mats = []
    for k in range(2):
        if uniform is None:
            alb = smooth_noise_texture(res, 3, seed * 10 + k, 0.1, 0.9)
            rough = smooth_noise_texture(res, 1, seed * 10 + k + 5, 0.2, 0.9)
        else:
            alb = np.full((res, res, 3), uniform, np.float32)
            rough = np.full((res, res), uniform, np.float32)
        if diffuse:
            mats.append(DiffuseMaterial(DiffuseMaterial.Parameters(BaseColor=TextureRgb(alb))))
            continue
        mats.append(GenericMaterial(GenericMaterial.Parameters(
            BaseColor=TextureRgb(alb), Roughness=TextureMono(rough), Metallic=0.0,
            SpecularTintStrength=spec_tint, SpecularTransmittance=0.0, IndexOfRefraction=1.5)))
    return mats[0], mats[1]


def large_scene(width: int = 1024, height: int = 1024, num_spheres: int = 400, grid: int = 512,
                scene_seed: int = 1, with_env: bool = True, camera_pos=(0.0, 3.0, 9.0),
                camera_target=(0.0, 0.5, 0.0), ground_textures: int = 0, ground_texture_seed: int = 5,
                ground_uniform: float | None = None) -> Scene:
    """The 1 036 296-triangle scene (sceneSeed 1) with the default arguments.

    ground_textures > 0 gives the ground grid a GenericMaterial with a
    ground_textures^2 RGB albedo and a mono roughness texture (UV-mapped over the
    grid): smooth noise from ground_texture_seed, or the constant ground_uniform."""
    s = Scene()
    s.Name = "large"
    sv, sf = icosphere(3)  # 1280 faces
    for i in range(num_spheres):
        r = Rng.seeded(scene_seed, i, 0)
        radius = r.next_float(0.02, 0.3)
        cx = r.next_float(-5.0, 5.0)
        cy = r.next_float(0.0, 4.0)
        cz = r.next_float(-5.0, 5.0)
        base = (r.next_float(), r.next_float(), r.next_float())
        rough = r.next_float(0.05, 1.0)
        metallic = 1.0 if r.next_float() < 0.3 else 0.0
        trans = 0.9 if r.next_float() < 0.1 else 0.0
        v = (sv * radius + np.array([cx, cy, cz])).astype(np.float32)
        m = Mesh(v, sf.reshape(-1), sv.astype(np.float32))
        m.Material = GenericMaterial(GenericMaterial.Parameters(
            BaseColor=base, Roughness=rough, Metallic=metallic, SpecularTintStrength=1.0,
            Anisotropic=0.0, SpecularTransmittance=trans, IndexOfRefraction=1.5))
        s.Meshes.append(m)
    r = Rng.seeded(scene_seed, num_spheres, 0)
    gp, gi, gn, guv = ground_grid(grid, 5.0, 0.2, scene_seed)
    ground = Mesh(gp, gi, gn, guv)
    ground.Material = GenericMaterial(GenericMaterial.Parameters(
        BaseColor=(r.next_float(), r.next_float(), r.next_float()), Roughness=r.next_float(0.05, 1.0),
        Metallic=0.0, SpecularTintStrength=1.0, SpecularTransmittance=0.0, IndexOfRefraction=1.5))
    if ground_textures:
        n = int(ground_textures)
        if ground_uniform is None:
            alb = smooth_noise_texture(n, 3, ground_texture_seed, 0.05, 0.95)
            rough = smooth_noise_texture(n, 1, ground_texture_seed + 50, 0.1, 0.9)
        else:
            alb = np.full((n, n, 3), ground_uniform, np.float32)
            rough = np.full((n, n), ground_uniform, np.float32)
        p = ground.Material.MaterialParameters
        p.BaseColor = TextureRgb(alb)
        p.Roughness = TextureMono(rough)
    s.Meshes.append(ground)
    black = DiffuseMaterial(DiffuseMaterial.Parameters(BaseColor=(0.0, 0.0, 0.0)))
    for (lx, lz) in ((-2.5, -2.5), (2.5, -2.5), (-2.5, 2.5), (2.5, 2.5)):
        lm = Mesh(*light_quad(lx, 6.0, lz, 1.0))
        lm.Material = black
        s.Meshes.append(lm)
        s.Emitters.extend(DiffuseEmitter.MakeFromMesh(lm, (20.0, 20.0, 20.0)))
    if with_env:
        s.Background = EnvironmentMap(sky_image())
    s.Camera = PerspectiveCamera(create_look_at(camera_pos, camera_target, (0, 1, 0)), 40.0)
    s.FrameBuffer = FrameBuffer(width, height)
    return s


def circle_views(count: int = 32, radius: float = 8.0, height: float = 2.0, target=(0.0, 0.5, 0.0),
                 fov: float = 40.0):
    """`count` PerspectiveCameras evenly spaced on a horizontal circle, looking at `target`."""
    cams = []
    for k in range(count):
        a = 2.0 * np.pi * k / count
        pos = (radius * np.cos(a), height, radius * np.sin(a))
        cams.append(PerspectiveCamera(create_look_at(pos, target, (0, 1, 0)), fov))
    return cams


BATCH_RESOLUTIONS = (128, 192, 256, 384, 512)
BATCH_SPP = (4, 16, 64)


def batch_scenes(count: int = 64, seed: int = 1000):
    """The batching workload: `count` independent scenes built like the large
    scene (spheres + fBm ground + 4 lights + sky), scene i seeded 1000 + i, with
    a triangle budget drawn log-uniformly from [1k, 200k] (half spheres, half
    ground grid), a resolution from {128..512}^2, a sample count from {4, 16, 64}
    and a random look-at camera. Returns (scenes, spp list)."""
    out, spps = [], []
    for i in range(count):
        r = Rng.seeded(seed + i, 0, 0)
        budget = math.exp(r.next_float(math.log(1000.0), math.log(200000.0)))
        num_spheres = max(1, int(budget / 2 / 1280))
        grid = max(8, int(math.sqrt(budget / 2 / 2)))
        res = BATCH_RESOLUTIONS[min(4, int(r.next_float() * 5))]
        spp = BATCH_SPP[min(2, int(r.next_float() * 3))]
        ang = r.next_float(0.0, 2 * math.pi)
        rad = r.next_float(6.0, 10.0)
        height = r.next_float(1.0, 4.0)
        target = (r.next_float(-1.0, 1.0), r.next_float(0.0, 1.0), r.next_float(-1.0, 1.0))
        sc = large_scene(res, res, num_spheres=num_spheres, grid=grid, scene_seed=seed + i,
                         camera_pos=(rad * math.cos(ang), height, rad * math.sin(ang)), camera_target=target)
        sc.Name = f"batch{i}"
        out.append(sc)
        spps.append(spp)
    return out, spps
