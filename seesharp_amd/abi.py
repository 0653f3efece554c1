"""ctypes mirror of include/sspt.h and the loader of libsspt.so.

The loader fails loudly when the native library is missing: there is no
Python or CPU fallback for any GPU entry point.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
# SSPT_LIBRARY selects another build of the library (e.g. one compiled with
# different occupancy targets); the default is the in-tree build.
LIB_PATH = os.environ.get("SSPT_LIBRARY") or os.path.join(_HERE, "libsspt.so")

ABI_VERSION = 8  # SSPT_ABI_VERSION of include/sspt.h
NOALLOC_SCRATCH_BYTES = 64 << 20  # SSPT_NOALLOC_SCRATCH_BYTES
SSPT_OK = 0
INVALID_ARGUMENT = 1
STATUS_NAMES = {
    0: "SSPT_OK",
    1: "SSPT_ERR_INVALID_ARGUMENT",
    2: "SSPT_ERR_INVALID_SCENE",
    3: "SSPT_ERR_DEVICE",
    4: "SSPT_ERR_OUT_OF_MEMORY",
    5: "SSPT_ERR_UNSUPPORTED",
    6: "SSPT_ERR_ZERO_AREA_EMITTER",
}
MAX_OUTLIERS = 4
EMITTER_DIFFUSE = 0
EMITTER_GLOSSY = 1
LAYER_ALBEDO = 0
LAYER_NORMAL = 1
EXEC_WAVEFRONT = 0
EXEC_SINGLE_KERNEL = 1
BVH_DEVICE = 0
BVH_HOST = 1
MATERIAL_DIFFUSE = 0
MATERIAL_GENERIC = 1

_fp = C.POINTER(C.c_float)
_ip = C.POINTER(C.c_int32)


class MeshDesc(C.Structure):
    _fields_ = [
        ("num_vertices", C.c_uint32),
        ("positions", _fp),
        ("normals", _fp),
        ("uvs", _fp),
        ("num_triangles", C.c_uint32),
        ("indices", _ip),
        ("material", C.c_int32),
    ]


class MaterialDesc(C.Structure):
    _fields_ = [
        ("kind", C.c_int32),
        ("base_color", C.c_float * 3),
        ("base_color_texture", C.c_int32),
        ("roughness", C.c_float),
        ("roughness_texture", C.c_int32),
        ("metallic", C.c_float),
        ("specular_tint_strength", C.c_float),
        ("anisotropic", C.c_float),
        ("specular_transmittance", C.c_float),
        ("index_of_refraction", C.c_float),
        ("transmitter", C.c_int32),
    ]


class TextureDesc(C.Structure):
    _fields_ = [
        ("width", C.c_uint32),
        ("height", C.c_uint32),
        ("channels", C.c_uint32),
        ("border", C.c_int32),
        ("texels", _fp),
    ]


class EmitterDesc(C.Structure):
    _fields_ = [("mesh", C.c_uint32), ("face", C.c_uint32), ("radiance", C.c_float * 3), ("kind", C.c_int32),
                ("exponent", C.c_float)]


class EnvmapDesc(C.Structure):
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32), ("rgb", _fp)]


class CameraDesc(C.Structure):
    _fields_ = [("world_to_camera", C.c_float * 16), ("vertical_fov_degrees", C.c_float)]


class SceneDesc(C.Structure):
    _fields_ = [
        ("num_meshes", C.c_uint32),
        ("meshes", C.POINTER(MeshDesc)),
        ("num_materials", C.c_uint32),
        ("materials", C.POINTER(MaterialDesc)),
        ("num_textures", C.c_uint32),
        ("textures", C.POINTER(TextureDesc)),
        ("num_emitters", C.c_uint32),
        ("emitters", C.POINTER(EmitterDesc)),
        ("background", C.POINTER(EnvmapDesc)),
        ("camera", CameraDesc),
        ("width", C.c_uint32),
        ("height", C.c_uint32),
    ]


class PtSettings(C.Structure):
    _fields_ = [
        ("base_seed", C.c_uint32),
        ("total_spp", C.c_int32),
        ("max_depth", C.c_int32),
        ("min_depth", C.c_int32),
        ("num_shadow_rays", C.c_int32),
        ("enable_bsdf_di", C.c_int32),
        ("max_render_time_ms", C.c_int64),
        ("enable_denoiser", C.c_int32),
    ]


class LtSettings(C.Structure):
    """sspt_lt_settings: the light tracer's BidirBase / Integrator fields."""
    _fields_ = [
        ("base_seed_light", C.c_uint32),
        ("num_light_paths", C.c_int32),
        ("max_depth", C.c_int32),
        ("min_depth", C.c_int32),
    ]


class BidirSettings(C.Structure):
    """sspt_bidir_settings: ClassicBidir's BidirBase / Integrator fields."""
    _fields_ = [
        ("base_seed_camera", C.c_uint32),
        ("base_seed_light", C.c_uint32),
        ("max_depth", C.c_int32),
        ("min_depth", C.c_int32),
        ("num_shadow_rays", C.c_int32),
        ("enable_connections", C.c_int32),
        ("enable_light_tracer", C.c_int32),
        ("num_light_paths", C.c_int32),
    ]


VERTEX_CACHE_BIDIR = 0  # sspt_vcm_algorithm
VCM = 1


class VcmSettings(C.Structure):
    """sspt_vcm_settings: VertexCacheBidir / VertexConnectionAndMerging fields."""
    _fields_ = [
        ("algorithm", C.c_int32),
        ("base_seed_camera", C.c_uint32),
        ("base_seed_light", C.c_uint32),
        ("max_depth", C.c_int32),
        ("min_depth", C.c_int32),
        ("num_light_paths", C.c_int32),
        ("num_connections", C.c_int32),
        ("num_shadow_rays", C.c_int32),
        ("enable_light_tracer", C.c_int32),
        ("enable_hitting", C.c_int32),
        ("enable_merging", C.c_int32),
        ("merge_primary", C.c_int32),
        ("max_num_photons", C.c_int32),
        ("disable_correl_aware_mis", C.c_int32),
    ]


class RayStats(C.Structure):
    _fields_ = [
        ("num_rays", C.c_uint64),
        ("num_shadow_rays", C.c_uint64),
        ("num_occluded", C.c_uint64),
        ("num_ray_hits", C.c_uint64),
        ("num_shadow_rays_traced", C.c_uint64),
        ("num_iterations", C.c_uint64),
    ]

    def as_dict(self):
        return {k: int(getattr(self, k)) for k, _ in self._fields_}


class Hit(C.Structure):
    _fields_ = [("t", C.c_float), ("u", C.c_float), ("v", C.c_float), ("prim", C.c_int32)]


NODE_CAMERA = 0
NODE_BSDF_SAMPLE = 1
NODE_NEXT_EVENT = 2
NODE_BACKGROUND = 3
NODE_HIT_EMITTER = 1
NODE_TO_BACKGROUND = 2


class PathNode(C.Structure):
    """sspt_path_node: one node of a replayed path's PathGraph."""
    _fields_ = [("kind", C.c_int32), ("ancestor", C.c_int32), ("position", C.c_float * 3),
                ("emission", C.c_float * 3), ("mis_weight", C.c_float), ("scatter_weight", C.c_float * 3),
                ("survival_probability", C.c_float), ("pdf", C.c_float), ("bsdf_cos", C.c_float * 3),
                ("prefix_weight", C.c_float * 3), ("flags", C.c_uint32)]


# Every symbol include/sspt.h declares, with (restype, argtypes).
_vp = C.c_void_p
SYMBOLS = {
    "sspt_abi_version": (C.c_int32, []),
    "sspt_ctx_create": (C.c_int, [C.c_int32, C.POINTER(_vp)]),
    "sspt_ctx_destroy": (None, [_vp]),
    "sspt_last_error": (C.c_char_p, [_vp]),
    "sspt_scene_create": (C.c_int, [_vp, C.POINTER(SceneDesc), C.POINTER(_vp)]),
    "sspt_scene_destroy": (None, [_vp]),
    "sspt_scene_info": (C.c_int, [_vp] + [C.POINTER(C.c_uint64)] * 4),
    "sspt_framebuffer_clear": (C.c_int, [_vp, _vp]),
    "sspt_framebuffer_read": (C.c_int, [_vp, _vp, _fp]),
    "sspt_framebuffer_nan_count": (C.c_int, [_vp, _vp, C.POINTER(C.c_uint64)]),
    "sspt_render": (C.c_int, [_vp, _vp, C.POINTER(PtSettings), C.c_uint32, C.c_uint32,
                              C.POINTER(RayStats)]),
    "sspt_render_noalloc": (C.c_int, [_vp, _vp, C.POINTER(PtSettings), C.c_uint32, C.c_uint32,
                                      C.POINTER(RayStats)]),
    "sspt_trace_closest": (C.c_int, [_vp, _vp, _fp, C.c_uint64, C.POINTER(Hit)]),
    "sspt_trace_occluded": (C.c_int, [_vp, _vp, _fp, C.c_uint64, C.POINTER(C.c_uint8)]),
    "sspt_primary_hits": (C.c_int, [_vp, _vp, C.c_uint32, C.c_uint32, _fp, C.POINTER(Hit)]),
    "sspt_batch_create": (C.c_int, [_vp, C.POINTER(SceneDesc), C.c_uint32, C.POINTER(_vp)]),
    "sspt_scene_count": (C.c_int, [_vp, C.POINTER(C.c_uint32)]),
    "sspt_render_batch": (C.c_int, [_vp, _vp, C.POINTER(PtSettings), C.POINTER(RayStats)]),
    "sspt_framebuffer_read_scene": (C.c_int, [_vp, _vp, C.c_uint32, _fp]),
    "sspt_render_adjoint": (C.c_int, [_vp, _vp, C.POINTER(PtSettings), _fp, C.c_uint32, C.c_uint32,
                                      C.POINTER(RayStats)]),
    "sspt_render_adjoint_batch": (C.c_int, [_vp, _vp, C.POINTER(PtSettings), _fp, C.c_uint32,
                                            C.POINTER(RayStats)]),
    "sspt_views_create": (C.c_int, [_vp, C.POINTER(SceneDesc), C.POINTER(CameraDesc), C.c_uint32,
                                    C.POINTER(_vp)]),
    "sspt_gradient_clear": (C.c_int, [_vp, _vp]),
    "sspt_gradient_read": (C.c_int, [_vp, _vp, C.c_uint32, _fp]),
    "sspt_texture_write": (C.c_int, [_vp, _vp, C.c_uint32, _fp]),
    "sspt_framebuffer_read_layer": (C.c_int, [_vp, _vp, C.c_uint32, C.c_int32, _fp]),
    "sspt_set_wave_capacity": (C.c_int, [_vp, C.c_uint32]),
    "sspt_set_execution_mode": (C.c_int, [_vp, C.c_int32]),
    "sspt_get_launch_config": (C.c_int, [_vp, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32),
                                         C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]),
    "sspt_set_launch_config": (C.c_int, [_vp, C.c_uint32, C.c_uint32, C.c_int32]),
    "sspt_set_block_shape": (C.c_int, [_vp, C.c_uint32, C.c_uint32]),
    "sspt_get_block_shape": (C.c_int, [_vp, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]),
    "sspt_render_light_tracer": (C.c_int, [_vp, _vp, C.POINTER(LtSettings), C.c_uint32, C.c_uint32,
                                           C.POINTER(RayStats)]),
    "sspt_render_bidir": (C.c_int, [_vp, _vp, C.POINTER(BidirSettings), C.c_uint32, C.c_uint32,
                                    C.POINTER(RayStats)]),
    "sspt_render_vcm": (C.c_int, [_vp, _vp, C.POINTER(VcmSettings), C.c_uint32, C.c_uint32,
                                  C.POINTER(RayStats)]),
    "sspt_set_bvh_builder": (C.c_int, [_vp, C.c_int32]),
    "sspt_set_bvh_leaf_size": (C.c_int, [_vp, C.c_int32]),
    "sspt_get_bvh_leaf_size": (C.c_int, [_vp, C.POINTER(C.c_int32)]),
    "sspt_get_bvh_builder": (C.c_int, [_vp, C.POINTER(C.c_int32)]),
    "sspt_scene_bvh_build_time": (C.c_int, [_vp, C.POINTER(C.c_int32), C.POINTER(C.c_double),
                                            C.POINTER(C.c_double)]),
    "sspt_scene_read_bvh": (C.c_int, [_vp, _vp, _fp, C.c_uint64, _ip, C.c_uint64]),
    "sspt_set_kernel_timing": (C.c_int, [_vp, C.c_int32]),
    "sspt_get_kernel_timing": (C.c_int, [_vp, C.c_char_p, C.POINTER(C.c_uint64),
                                         C.POINTER(C.c_double)]),
    "sspt_reset_kernel_timing": (C.c_int, [_vp]),
    "sspt_framebuffer_track_outliers": (C.c_int, [_vp, _vp, C.c_uint32]),
    "sspt_framebuffer_read_outliers": (C.c_int, [_vp, _vp, C.c_uint32, _fp, C.POINTER(C.c_int32)]),
    "sspt_render_debug": (C.c_int, [_vp, _vp, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32]),
    "sspt_replay_pixel": (C.c_int, [_vp, _vp, C.c_uint32, C.POINTER(PtSettings), C.c_uint32, C.c_uint32,
                                    C.c_uint32, C.POINTER(PathNode), C.c_uint32, C.POINTER(C.c_uint32), _fp]),
}

_lib = None


class NativeLibraryMissing(RuntimeError):
    pass


def load_library(path: str | None = None):
    """Loads libsspt.so once and binds every exported entry point."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise NativeLibraryMissing(
            f"libsspt.so not found at {p}: build it with `python -c 'import __graft_entry__ as g; g.build()'`"
        )
    lib = C.CDLL(p)
    for name, (res, args) in SYMBOLS.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if path is None:
        _lib = lib
    return lib


class SsptError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__(f"{STATUS_NAMES.get(status, status)}: {message}")
        self.status = status
        self.message = message


def fptr(a):
    return a.ctypes.data_as(_fp) if a is not None else None


def iptr(a):
    return a.ctypes.data_as(_ip) if a is not None else None
