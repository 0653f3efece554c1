"""ctypes binding of the CPU oracle (oracle/liboracle.so) for tests and bench.py's CPU baseline.

The oracle is the checker: tests compare the GPU library against it; it is never
the thing measured as the product or shipped.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

from seesharp_amd import abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
ORACLE_SO = os.path.join(ORACLE_DIR, "liboracle.so")

_fp = C.POINTER(C.c_float)
_u32p = C.POINTER(C.c_uint32)
_vp = C.c_void_p

_SIGS = {
    "or_hash_seed": (C.c_uint32, [C.c_uint32] * 3),
    "or_pcg_hash": (C.c_uint32, [C.c_uint32]),
    "or_fnv_hash": (C.c_uint32, [C.c_uint32, C.c_uint32]),
    "or_rng_stream": (None, [C.c_uint32, C.c_uint32, _u32p, _fp]),
    "or_rng_next_int": (None, [C.c_uint32, C.c_uint32, C.c_uint32, _u32p]),
    "or_warp_uniform_triangle": (None, [_fp, _fp]),
    "or_warp_cos_hemisphere": (None, [_fp, _fp, _fp]),
    "or_pwc_sample": (None, [_fp, C.c_int32, C.c_float, C.POINTER(C.c_int32), _fp]),
    "or_pwc_probability": (C.c_float, [_fp, C.c_int32, C.c_int32]),
    "or_camera_ray": (None, [_fp, C.c_float, C.c_int32, C.c_int32, C.c_float, C.c_float, _fp]),
    "or_look_at": (None, [_fp, _fp, _fp, _fp]),
    "or_ggx_ndf": (C.c_float, [C.c_float, C.c_float, _fp]),
    "or_ggx_pdf": (C.c_float, [C.c_float, C.c_float, _fp, _fp]),
    "or_ggx_sample": (None, [C.c_float, C.c_float, _fp, _fp, _fp]),
    "or_fresnel_dielectric": (C.c_float, [C.c_float] * 3),
    "or_material_eval": (None, [C.POINTER(abi.MaterialDesc), C.POINTER(abi.TextureDesc), _fp, _fp, _fp,
                                _fp, _fp, _fp, _fp]),
    "or_material_sample": (None, [C.POINTER(abi.MaterialDesc), C.POINTER(abi.TextureDesc), _fp, _fp, _fp,
                                  C.c_float, _fp, _fp]),
    "or_material_pdf_pair": (None, [C.POINTER(abi.MaterialDesc), C.POINTER(abi.TextureDesc), _fp, _fp, _fp,
                                    _fp, _fp, _fp]),
    "or_material_eval_many": (None, [C.POINTER(abi.MaterialDesc), C.POINTER(abi.TextureDesc), _fp, _fp, _fp,
                                     C.c_int64, _fp, _fp, _fp]),
    "or_material_integrate": (None, [C.POINTER(abi.MaterialDesc), C.POINTER(abi.TextureDesc), _fp, _fp, _fp,
                                     C.c_int64, C.c_uint32, C.c_int32, C.POINTER(C.c_double),
                                     C.POINTER(C.c_double)]),
    "or_scene_create": (_vp, [C.POINTER(abi.SceneDesc), C.c_char_p]),
    "or_scene_destroy": (None, [_vp]),
    "or_env_eval": (None, [_vp, _fp, _fp]),
    "or_env_pdf": (C.c_float, [_vp, _fp]),
    "or_env_sample": (None, [_vp, _fp, _fp]),
    "or_trace_closest": (None, [_vp, _fp, C.c_uint64, C.POINTER(abi.Hit), C.c_int32]),
    "or_trace_occluded": (None, [_vp, _fp, C.c_uint64, C.POINTER(C.c_uint8), C.c_int32]),
    "or_primary_hits": (None, [_vp, C.c_uint32, C.c_uint32, _fp, C.POINTER(abi.Hit), C.c_int32]),
    "or_render": (C.c_uint64, [_vp, C.POINTER(abi.PtSettings), C.c_uint32, C.c_uint32, _fp,
                               C.POINTER(abi.RayStats), C.POINTER(C.c_int32), C.c_int64, C.c_int32]),
    "or_render_light_tracer": (C.c_uint64, [_vp, C.POINTER(abi.LtSettings), C.c_uint32, C.c_uint32, _fp,
                                            C.POINTER(abi.RayStats), C.c_int32]),
    "or_render_bidir": (C.c_uint64, [_vp, C.POINTER(abi.BidirSettings), C.c_uint32, C.c_uint32, _fp,
                                     C.POINTER(abi.RayStats), C.c_int32]),
    "or_render_vcm": (C.c_uint64, [_vp, C.POINTER(abi.VcmSettings), C.c_float, C.c_uint32, C.c_uint32, _fp,
                                   C.POINTER(abi.RayStats), C.c_int32]),
    "or_render_pixel": (None, [_vp, C.POINTER(abi.PtSettings), C.c_uint32, C.c_uint32, C.c_uint32, _fp]),
    "or_set_texture": (C.c_int32, [_vp, C.c_uint32, _fp]),
    "or_render_outliers": (None, [_vp, C.POINTER(abi.PtSettings), C.c_uint32, C.c_uint32, C.c_int32, _fp,
                                  C.POINTER(C.c_int32), C.c_int32]),
    "or_render_debug": (None, [_vp, C.c_uint32, C.c_uint32, C.c_uint32, _fp, C.c_int32]),
    "or_replay_pixel": (C.c_uint32, [_vp, C.POINTER(abi.PtSettings), C.c_uint32, C.c_uint32, C.c_uint32,
                                     C.POINTER(abi.PathNode), C.c_uint32, _fp]),
    "or_primary_aov": (None, [_vp, C.POINTER(abi.PtSettings), C.c_uint32, C.c_uint32, _fp, _fp, C.c_int32]),
    "or_render_tangent": (C.c_uint64, [_vp, C.POINTER(abi.PtSettings), C.c_uint32, C.c_uint32, C.c_uint32, _fp,
                                       C.c_float, _fp, C.POINTER(C.c_int32), C.c_int64, C.c_int32]),
}

_lib = None


def build():
    subprocess.run(["make", "-s", "-C", ORACLE_DIR], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_SO):
            build()
        L = C.CDLL(ORACLE_SO)
        for k, (r, a) in _SIGS.items():
            f = getattr(L, k)
            f.restype = r
            f.argtypes = a
        _lib = L
    return _lib


def fp(a):
    return a.ctypes.data_as(_fp)


def lt_settings(base_seed_light=0x13C0FEFE, num_light_paths=-1, max_depth=5, min_depth=1):
    # (MaxDepth defaults to 100 in the reference's Integrator; the tests pass small depths)
    return abi.LtSettings(base_seed_light, num_light_paths, max_depth, min_depth)


def bidir_settings(base_seed_camera=0xC030114, base_seed_light=0x13C0FEFE, max_depth=5, min_depth=1,
                   num_shadow_rays=1, enable_connections=True, enable_light_tracer=True, num_light_paths=-1):
    return abi.BidirSettings(base_seed_camera, base_seed_light, max_depth, min_depth, num_shadow_rays,
                             1 if enable_connections else 0, 1 if enable_light_tracer else 0, num_light_paths)


def vcm_settings(algorithm=abi.VCM, base_seed_camera=0xC030114, base_seed_light=0x13C0FEFE, max_depth=5,
                 min_depth=1, num_light_paths=-1, num_connections=1, num_shadow_rays=1, enable_light_tracer=True,
                 enable_hitting=True, enable_merging=True, merge_primary=False, max_num_photons=8,
                 disable_correl_aware_mis=False):
    return abi.VcmSettings(algorithm, base_seed_camera, base_seed_light, max_depth, min_depth, num_light_paths,
                           num_connections, num_shadow_rays, int(enable_light_tracer), int(enable_hitting),
                           int(enable_merging), int(merge_primary), max_num_photons, int(disable_correl_aware_mis))


def settings(base_seed=0xC030114, total_spp=1, max_depth=5, min_depth=1, num_shadow_rays=1,
             enable_bsdf_di=True, max_render_time_ms=-1, enable_denoiser=False):
    return abi.PtSettings(base_seed, total_spp, max_depth, min_depth, num_shadow_rays,
                          1 if enable_bsdf_di else 0, max_render_time_ms, 1 if enable_denoiser else 0)


class OracleScene:
    def __init__(self, scene):
        self.scene = scene
        self.desc = scene.to_desc()
        err = C.create_string_buffer(256)
        self.ptr = lib().or_scene_create(C.byref(self.desc), err)
        if not self.ptr:
            raise ValueError(err.value.decode())
        self.width = scene.FrameBuffer.Width
        self.height = scene.FrameBuffer.Height

    def close(self):
        if self.ptr:
            lib().or_scene_destroy(self.ptr)
            self.ptr = None

    def __del__(self):
        self.close()

    def render(self, cfg, first=0, num=1, image=None, pixels=None, threads=0):
        if image is None:
            image = np.zeros((self.height, self.width, 3), np.float32)
        stats = abi.RayStats()
        px = None
        n = 0
        if pixels is not None:
            pixels = np.ascontiguousarray(pixels, np.int32)
            px = pixels.ctypes.data_as(C.POINTER(C.c_int32))
            n = pixels.size
        nans = lib().or_render(self.ptr, C.byref(cfg), first, num, fp(image), C.byref(stats), px, n, threads)
        return image, stats, int(nans)

    def render_light_tracer(self, cfg, first=0, num=1, image=None, threads=0):
        """The light tracer's running-mean image after iterations first..first+num-1,
        its ray statistics and the non-finite splat count."""
        if image is None:
            image = np.zeros((self.height, self.width, 3), np.float32)
        stats = abi.RayStats()
        nans = lib().or_render_light_tracer(self.ptr, C.byref(cfg), first, num, fp(image), C.byref(stats), threads)
        if nans == 2 ** 64 - 1:
            raise ValueError("the light tracer needs emitters or a background")
        return image, stats, int(nans)

    def render_bidir(self, cfg, first=0, num=1, image=None, threads=0):
        """ClassicBidir's running-mean image after iterations first..first+num-1,
        its ray statistics and the non-finite value count."""
        if image is None:
            image = np.zeros((self.height, self.width, 3), np.float32)
        stats = abi.RayStats()
        nans = lib().or_render_bidir(self.ptr, C.byref(cfg), first, num, fp(image), C.byref(stats), threads)
        if nans == 2 ** 64 - 1:
            raise ValueError("ClassicBidir needs emitters, no background and one light path per pixel")
        return image, stats, int(nans)

    def render_vcm(self, cfg, first=0, num=1, image=None, threads=0, cell=None):
        """VertexCacheBidir / VCM running-mean image after iterations first..first+num-1, its
        ray statistics and the non-finite value count. `cell`: the photon grid's cell size
        (default: the scene's bounding-box diagonal over twice the larger image side)."""
        if image is None:
            image = np.zeros((self.height, self.width, 3), np.float32)
        if cell is None:
            pts = np.concatenate([m.Vertices for m in self.scene.Meshes])
            cell = float(np.linalg.norm(pts.max(axis=0) - pts.min(axis=0))) / (2 * max(self.width, self.height))
        stats = abi.RayStats()
        nans = lib().or_render_vcm(self.ptr, C.byref(cfg), cell, first, num, fp(image), C.byref(stats), threads)
        if nans == 2 ** 64 - 1:
            raise ValueError("VertexCacheBidir / VCM need emitters, no background and valid settings")
        return image, stats, int(nans)

    def primary_aov(self, cfg, first=0, num=1, threads=0):
        """(albedo, normal) running-mean layers of iterations first..first+num-1."""
        alb = np.zeros((self.height, self.width, 3), np.float32)
        nrm = np.zeros((self.height, self.width, 3), np.float32)
        lib().or_primary_aov(self.ptr, C.byref(cfg), first, num, fp(alb), fp(nrm), threads)
        return alb, nrm

    def replay_pixel(self, cfg, col, row, iteration):
        """ReplayPixel: (sspt_path_node list, RGB estimate) of one path."""
        cap = 2 + cfg.max_depth * (1 + 2 * max(0, cfg.num_shadow_rays))
        nodes = (abi.PathNode * cap)()
        est = np.zeros(3, np.float32)
        n = lib().or_replay_pixel(self.ptr, C.byref(cfg), col, row, iteration, nodes, cap, fp(est))
        assert n <= cap
        return list(nodes[:n]), est

    def render_outliers(self, cfg, first, num, n, threads=0):
        """(weights, iterations), each H x W x n: the outlier cache after the render."""
        w = np.zeros((self.height, self.width, n), np.float32)
        it = np.zeros((self.height, self.width, n), np.int32)
        lib().or_render_outliers(self.ptr, C.byref(cfg), first, num, n, fp(w),
                                 it.ctypes.data_as(C.POINTER(C.c_int32)), threads)
        return w, it

    def render_debug(self, base_seed, first, num, threads=0):
        img = np.zeros((self.height, self.width, 3), np.float32)
        lib().or_render_debug(self.ptr, base_seed, first, num, fp(img), threads)
        return img

    def render_pixel(self, cfg, col, row, iteration):
        out = np.zeros(3, np.float32)
        lib().or_render_pixel(self.ptr, C.byref(cfg), row, col, iteration, fp(out))
        return out

    def set_texture(self, index, texels):
        texels = np.ascontiguousarray(texels, np.float32)
        if lib().or_set_texture(self.ptr, index, fp(texels)) != 0:
            raise IndexError(index)

    def render_tangent(self, cfg, first, num, texture_index, tangent, eps=1e-3, pixels=None, threads=0):
        """Per-pixel d(estimate)/dt for texels + t * tangent, samples held fixed."""
        tangent = np.ascontiguousarray(tangent, np.float32)
        out = np.zeros((self.height, self.width, 3), np.float32)
        px = None
        n = 0
        if pixels is not None:
            pixels = np.ascontiguousarray(pixels, np.int32)
            px = pixels.ctypes.data_as(C.POINTER(C.c_int32))
            n = pixels.size
        nans = lib().or_render_tangent(self.ptr, C.byref(cfg), first, num, texture_index, fp(tangent), eps,
                                       fp(out), px, n, threads)
        return out, int(nans)

    def trace_closest(self, rays, brute=False):
        rays = np.ascontiguousarray(rays, np.float32).reshape(-1, 8)
        hits = (abi.Hit * rays.shape[0])()
        lib().or_trace_closest(self.ptr, fp(rays), rays.shape[0], hits, 1 if brute else 0)
        return hits_to_numpy(hits)

    def trace_occluded(self, rays, brute=False):
        rays = np.ascontiguousarray(rays, np.float32).reshape(-1, 8)
        out = np.zeros(rays.shape[0], np.uint8)
        lib().or_trace_occluded(self.ptr, fp(rays), rays.shape[0],
                                out.ctypes.data_as(C.POINTER(C.c_uint8)), 1 if brute else 0)
        return out

    def primary_hits(self, base_seed, iteration, brute=False):
        n = self.width * self.height
        rays = np.zeros((n, 8), np.float32)
        hits = (abi.Hit * n)()
        lib().or_primary_hits(self.ptr, base_seed, iteration, fp(rays), hits, 1 if brute else 0)
        return rays, hits_to_numpy(hits)


def hits_to_numpy(hits):
    arr = np.frombuffer(hits, dtype=np.dtype([("t", "<f4"), ("u", "<f4"), ("v", "<f4"), ("prim", "<i4")]))
    return arr.copy()
