"""GpuPathTracer: the drop-in `Integrator` over libsspt.so.

Mirrors ``Integrator`` (SeeSharp/Integrators/Integrator.cs:6-30) and the public
settings of ``PathTracerBase`` (SeeSharp/Integrators/PathTracer.cs:16-49):
BaseSeed, TotalSpp, MaximumRenderTimeMs, NumShadowRays, EnableBsdfDI,
MaxDepth, MinDepth. ``Render(scene)`` runs every iteration on the GPU and
leaves the running mean in ``scene.FrameBuffer.Image`` with the same
metadata keys the reference writes (RenderTime, NumIterations,
RayTracerStats; PathTracer.cs:254-257, FrameBuffer.cs:169-180).

There is no CPU fallback: without libsspt.so or a GPU every call raises.
"""
from __future__ import annotations

import ctypes as C
import time

import numpy as np

from . import abi, metrics


class Device:
    """One HIP device context (sspt_ctx) with its own stream."""

    def __init__(self, device: int = 0):
        self.lib = abi.load_library()
        h = C.c_void_p()
        st = self.lib.sspt_ctx_create(device, C.byref(h))
        if st != abi.SSPT_OK:
            raise abi.SsptError(st, f"sspt_ctx_create(device={device}) failed (no usable HIP device?)")
        self.ptr = h

    def check(self, st: int):
        if st != abi.SSPT_OK:
            raise abi.SsptError(st, self.lib.sspt_last_error(self.ptr).decode())

    def close(self):
        if getattr(self, "ptr", None):
            self.lib.sspt_ctx_destroy(self.ptr)
            self.ptr = None

    def __del__(self):
        self.close()

    def set_wave_capacity(self, slots: int):
        self.check(self.lib.sspt_set_wave_capacity(self.ptr, int(slots)))

    def set_launch_config(self, trace_blocks: int = 0, shade_blocks: int = 0, use_graphs: bool = True):
        self.check(self.lib.sspt_set_launch_config(self.ptr, int(trace_blocks), int(shade_blocks),
                                                   1 if use_graphs else 0))

    def set_block_shape(self, trace_block: int = 0, shade_block: int = 0):
        """Threads per traversal block (64, 128, 256) and per forward shade block
        (128, 256, 512); 0 = the defaults."""
        self.check(self.lib.sspt_set_block_shape(self.ptr, int(trace_block), int(shade_block)))

    def block_shape(self) -> tuple[int, int]:
        t, s = C.c_uint32(), C.c_uint32()
        self.check(self.lib.sspt_get_block_shape(self.ptr, C.byref(t), C.byref(s)))
        return int(t.value), int(s.value)

    def set_bvh_builder(self, host: bool):
        """BVH builder of the scenes created from now on: the device (default) or the host."""
        self.check(self.lib.sspt_set_bvh_builder(self.ptr, abi.BVH_HOST if host else abi.BVH_DEVICE))

    def set_bvh_leaf_size(self, max_triangles: int):
        """Triangles per BVH leaf at most (1..16, default 2) for the scenes created from now on."""
        self.check(self.lib.sspt_set_bvh_leaf_size(self.ptr, int(max_triangles)))

    def bvh_leaf_size(self) -> int:
        n = C.c_int32()
        self.check(self.lib.sspt_get_bvh_leaf_size(self.ptr, C.byref(n)))
        return n.value

    def bvh_builder(self) -> str:
        b = C.c_int32()
        self.check(self.lib.sspt_get_bvh_builder(self.ptr, C.byref(b)))
        return "host" if b.value == abi.BVH_HOST else "device"

    def set_execution_mode(self, single_kernel: bool):
        """False: the wavefront tracer (default). True: the plain single-kernel tracer."""
        self.check(self.lib.sspt_set_execution_mode(self.ptr, abi.EXEC_SINGLE_KERNEL if single_kernel
                                                    else abi.EXEC_WAVEFRONT))

    def launch_config(self) -> dict:
        v = [C.c_uint32() for _ in range(4)]
        self.check(self.lib.sspt_get_launch_config(self.ptr, *[C.byref(x) for x in v]))
        return {"trace_blocks": v[0].value, "shade_blocks": v[1].value, "trace_per_cu": v[2].value,
                "shade_per_cu": v[3].value}

    # kernel timing with HIP events on the context stream
    def set_kernel_timing(self, on: bool):
        self.check(self.lib.sspt_set_kernel_timing(self.ptr, 1 if on else 0))

    def reset_kernel_timing(self):
        self.check(self.lib.sspt_reset_kernel_timing(self.ptr))

    def kernel_timing(self, name: str):
        n = C.c_uint64()
        ms = C.c_double()
        self.check(self.lib.sspt_get_kernel_timing(self.ptr, name.encode(), C.byref(n), C.byref(ms)))
        return int(n.value), float(ms.value)


def _image_pointer(image, size: int):
    """(float pointer, owner) of a float32 image: a torch tensor on the GPU by its device
    pointer (the library copies device to device), anything else as a host array."""
    if hasattr(image, "data_ptr") and getattr(image, "is_cuda", False):
        import torch

        if image.dtype != torch.float32 or not image.is_contiguous() or image.numel() != size:
            raise ValueError(f"device dLoss/dImage must be a contiguous float32 tensor of {size} values")
        # the library reads it on its own stream: torch's work on it must be done
        torch.cuda.current_stream(image.device).synchronize()
        return C.cast(C.c_void_p(image.data_ptr()), C.POINTER(C.c_float)), image
    arr = np.ascontiguousarray(image, np.float32).reshape(-1)
    if arr.size != size:
        raise ValueError(f"dLoss/dImage holds {arr.size} values, the framebuffer {size}")
    return abi.fptr(arr), arr


class _TextureParameters:
    """Texture upload and gradient read-back of a device world (the
    differentiable parameters). Textures are numbered in world order: the
    textures of each scene description, scene after scene."""

    def _texture_owners(self):
        raise NotImplementedError

    def texture_index(self, texture) -> int:
        """Device index of a TextureRgb / TextureMono object of the world's scenes."""
        base = 0
        for scene in self._texture_owners():
            for k, t in enumerate(scene.DeviceTextures):
                if t is texture:
                    return base + k
            base += len(scene.DeviceTextures)
        raise KeyError("texture is not part of this world (or has no image)")

    def texture_shapes(self):
        """(height, width[, channels]) of each texture, in world order."""
        shapes = []
        for scene in self._texture_owners():
            for t in scene.DeviceTextures:
                img = np.asarray(t.Image)
                shapes.append(tuple(img.shape))
        return shapes

    def write_texture(self, index: int, texels: np.ndarray):
        texels = np.ascontiguousarray(texels, np.float32)
        shapes = self.texture_shapes()
        if 0 <= index < len(shapes) and texels.size != int(np.prod(shapes[index])):
            raise ValueError(f"texture {index} holds {int(np.prod(shapes[index]))} texels, got {texels.size}")
        self.device.check(self.device.lib.sspt_texture_write(self.device.ptr, self.ptr, index, abi.fptr(texels)))

    def clear_gradients(self):
        self.device.check(self.device.lib.sspt_gradient_clear(self.device.ptr, self.ptr))

    def gradient(self, index: int) -> np.ndarray:
        out = np.zeros(self.texture_shapes()[index], np.float32)
        self.device.check(self.device.lib.sspt_gradient_read(self.device.ptr, self.ptr, index, abi.fptr(out)))
        return out


class GpuScene(_TextureParameters):
    """The device copy of a prepared Scene: BVH, triangles, materials, emitters,
    environment sampling grid and the framebuffer (sspt_scene)."""

    def __init__(self, device: Device, scene):
        self.device = device
        self.scene = scene
        self.desc = scene.to_desc()
        h = C.c_void_p()
        device.check(device.lib.sspt_scene_create(device.ptr, C.byref(self.desc), C.byref(h)))
        self.ptr = h
        self.width = scene.FrameBuffer.Width
        self.height = scene.FrameBuffer.Height
        self.num_outliers = 0

    def close(self):
        if getattr(self, "ptr", None):
            self.device.lib.sspt_scene_destroy(self.ptr)
            self.ptr = None

    def __del__(self):
        self.close()

    def info(self):
        v = [C.c_uint64() for _ in range(4)]
        self.device.check(self.device.lib.sspt_scene_info(self.ptr, *[C.byref(x) for x in v]))
        return dict(zip(("num_triangles", "num_emitters", "num_bvh_nodes", "device_bytes"), (x.value for x in v)))

    def bvh_build_time(self) -> dict:
        """Builder, summed build-kernel milliseconds (device builder) and wall milliseconds of the BVH build."""
        b, k, w = C.c_int32(), C.c_double(), C.c_double()
        self.device.check(self.device.lib.sspt_scene_bvh_build_time(self.ptr, C.byref(b), C.byref(k), C.byref(w)))
        return {"builder": "host" if b.value == abi.BVH_HOST else "device", "kernel_ms": k.value,
                "wall_ms": w.value}

    def read_bvh(self):
        """The 4-wide BVH as the builder made it (the traversal reads the same tree with
        fp16 child planes rounded outward): nodes (N, 32) float32 (child box planes, then child
        indices and counts as int32 bits) and the global triangle id of each leaf position."""
        info = self.info()
        nodes = np.zeros((info["num_bvh_nodes"], 32), np.float32)
        tris = np.zeros(info["num_triangles"], np.int32)
        self.device.check(self.device.lib.sspt_scene_read_bvh(
            self.device.ptr, self.ptr, abi.fptr(nodes), nodes.shape[0],
            tris.ctypes.data_as(C.POINTER(C.c_int32)), tris.size))
        return nodes, tris

    def clear(self):
        self.device.check(self.device.lib.sspt_framebuffer_clear(self.device.ptr, self.ptr))

    def read_image(self) -> np.ndarray:
        img = np.zeros((self.height, self.width, 3), np.float32)
        self.device.check(self.device.lib.sspt_framebuffer_read(self.device.ptr, self.ptr, abi.fptr(img)))
        return img

    def read_layer(self, name: str) -> np.ndarray:
        """The "albedo" or "normal" denoiser layer (H, W, 3) of renders with enable_denoiser."""
        return _read_layer(self.device, self.ptr, 0, name, self.height, self.width)

    def nan_count(self) -> int:
        n = C.c_uint64()
        self.device.check(self.device.lib.sspt_framebuffer_nan_count(self.device.ptr, self.ptr, C.byref(n)))
        return int(n.value)

    def render(self, settings: abi.PtSettings, first_iteration: int, num_iterations: int) -> dict:
        stats = abi.RayStats()
        self.device.check(self.device.lib.sspt_render(self.device.ptr, self.ptr, C.byref(settings),
                                                      first_iteration, num_iterations, C.byref(stats)))
        return stats.as_dict()

    def render_noalloc(self, settings: abi.PtSettings, first_iteration: int, num_iterations: int) -> dict:
        """render() without device allocation: the wave buffers live in the
        library's per-GPU scratch array (sspt_render_noalloc)."""
        stats = abi.RayStats()
        self.device.check(self.device.lib.sspt_render_noalloc(self.device.ptr, self.ptr, C.byref(settings),
                                                              first_iteration, num_iterations, C.byref(stats)))
        return stats.as_dict()

    def render_mean(self, settings: abi.PtSettings, first_iteration: int, num_iterations: int) -> np.ndarray:
        """Mean of the estimates of iterations first..first+num-1 as an (H, W, 3)
        image. The framebuffer follows FrameBuffer's running mean, which after a
        clear and iterations starting at `first` holds their sum / (first + num);
        this rescales it to the plain mean."""
        self.clear()
        self.render(settings, first_iteration, num_iterations)
        img = self.read_image()
        if first_iteration:
            img *= np.float32((first_iteration + num_iterations) / num_iterations)
        return img

    # ---- differentiation (path-replay adjoint)
    def _texture_owners(self):
        return [self.scene]

    def render_adjoint(self, settings: abi.PtSettings, dloss_dimage, first_iteration: int,
                       num_iterations: int) -> dict:
        """Adds dLoss/dtexel for the image of iterations first..first+num-1 to the gradients.
        dloss_dimage: an (H, W, 3) array, or a contiguous float32 torch tensor on the GPU
        (passed by its device pointer)."""
        ptr, keep = _image_pointer(dloss_dimage, self.height * self.width * 3)
        stats = abi.RayStats()
        self.device.check(self.device.lib.sspt_render_adjoint(self.device.ptr, self.ptr, C.byref(settings),
                                                              ptr, first_iteration, num_iterations,
                                                              C.byref(stats)))
        del keep
        return stats.as_dict()

    def trace_closest(self, rays: np.ndarray) -> np.ndarray:
        rays = np.ascontiguousarray(rays, np.float32).reshape(-1, 8)
        hits = (abi.Hit * rays.shape[0])()
        self.device.check(self.device.lib.sspt_trace_closest(self.device.ptr, self.ptr, abi.fptr(rays),
                                                             rays.shape[0], hits))
        return _hits(hits)

    def trace_occluded(self, rays: np.ndarray) -> np.ndarray:
        rays = np.ascontiguousarray(rays, np.float32).reshape(-1, 8)
        out = np.zeros(rays.shape[0], np.uint8)
        self.device.check(self.device.lib.sspt_trace_occluded(
            self.device.ptr, self.ptr, abi.fptr(rays), rays.shape[0], out.ctypes.data_as(C.POINTER(C.c_uint8))))
        return out

    def render_light_tracer(self, settings: abi.LtSettings, first_iteration: int, num_iterations: int) -> dict:
        """Iterations first..first+num-1 of the light tracer into the framebuffer (running mean)."""
        stats = abi.RayStats()
        self.device.check(self.device.lib.sspt_render_light_tracer(
            self.device.ptr, self.ptr, C.byref(settings), first_iteration, num_iterations, C.byref(stats)))
        return stats.as_dict()

    def render_bidir(self, settings: abi.BidirSettings, first_iteration: int, num_iterations: int) -> dict:
        """Iterations first..first+num-1 of ClassicBidir into the framebuffer (running mean)."""
        stats = abi.RayStats()
        self.device.check(self.device.lib.sspt_render_bidir(
            self.device.ptr, self.ptr, C.byref(settings), first_iteration, num_iterations, C.byref(stats)))
        return stats.as_dict()

    def render_vcm(self, settings: abi.VcmSettings, first_iteration: int, num_iterations: int) -> dict:
        """Iterations first..first+num-1 of VertexCacheBidir / VCM into the framebuffer (running mean)."""
        stats = abi.RayStats()
        self.device.check(self.device.lib.sspt_render_vcm(
            self.device.ptr, self.ptr, C.byref(settings), first_iteration, num_iterations, C.byref(stats)))
        return stats.as_dict()

    def render_debug(self, base_seed: int, first_iteration: int, num_iterations: int):
        """DebugVisualizer iterations into the running-mean framebuffer."""
        self.device.check(self.device.lib.sspt_render_debug(self.device.ptr, self.ptr, 0, base_seed & 0xFFFFFFFF,
                                                            first_iteration, num_iterations))

    def track_outliers(self, n: int):
        """Outlier tracking of n <= 4 samples per pixel (0 = off), starting empty."""
        self.device.check(self.device.lib.sspt_framebuffer_track_outliers(self.device.ptr, self.ptr, n))
        self.num_outliers = n

    def read_outliers(self):
        """(weights, iterations), each H x W x n, largest first."""
        return _read_outliers(self.device, self.ptr, 0, self.height, self.width, self.num_outliers)

    def replay_pixel(self, settings: abi.PtSettings, col: int, row: int, iteration: int):
        """ReplayPixel (PathTracer.cs:61-69): the sspt_path_node records and the
        RGB estimate of the path of pixel (col, row) in `iteration`."""
        return _replay(self.device, self.ptr, 0, settings, col, row, iteration)

    def primary_hits(self, base_seed: int, iteration: int):
        n = self.width * self.height
        rays = np.zeros((n, 8), np.float32)
        hits = (abi.Hit * n)()
        self.device.check(self.device.lib.sspt_primary_hits(self.device.ptr, self.ptr, base_seed, iteration,
                                                            abi.fptr(rays), hits))
        return rays, _hits(hits)


class GpuBatch(_TextureParameters):
    """Many scenes in one device world, rendered together (sspt_batch_create /
    sspt_render_batch): waves are filled with paths of every scene, so small
    scenes keep the GPU busy. `GpuBatch.views` builds the same for many cameras
    on one scene (sspt_views_create): one copy of the geometry and textures,
    one framebuffer per view."""

    def __init__(self, device: Device, scenes, _views=None):
        self.device = device
        self.scenes = list(scenes)
        descs = [s.to_desc() for s in self.scenes]
        self._descs = (abi.SceneDesc * len(descs))(*descs)
        h = C.c_void_p()
        if _views is None:
            self._owners = self.scenes
            self.sizes = [(s.FrameBuffer.Height, s.FrameBuffer.Width) for s in self.scenes]
            device.check(device.lib.sspt_batch_create(device.ptr, self._descs, len(descs), C.byref(h)))
        else:
            cams = (abi.CameraDesc * len(_views))()
            for i, cam in enumerate(_views):
                cams[i].world_to_camera[:] = [float(x) for x in cam.WorldToCamera.reshape(16)]
                cams[i].vertical_fov_degrees = cam.VerticalFieldOfView
            self._cams = cams
            self._owners = self.scenes[:1]
            fb = self.scenes[0].FrameBuffer
            self.sizes = [(fb.Height, fb.Width)] * len(_views)
            device.check(device.lib.sspt_views_create(device.ptr, self._descs, cams, len(_views), C.byref(h)))
        self.ptr = h

    @classmethod
    def views(cls, device: Device, scene, cameras) -> "GpuBatch":
        """One scene seen from each PerspectiveCamera in `cameras` (the scene's
        own camera is not used), at the scene's framebuffer size."""
        return cls(device, [scene], _views=list(cameras))

    def __len__(self):
        return len(self.sizes)

    def _texture_owners(self):
        return self._owners

    def close(self):
        if getattr(self, "ptr", None):
            self.device.lib.sspt_scene_destroy(self.ptr)
            self.ptr = None

    def __del__(self):
        self.close()

    def info(self):
        v = [C.c_uint64() for _ in range(4)]
        self.device.check(self.device.lib.sspt_scene_info(self.ptr, *[C.byref(x) for x in v]))
        return dict(zip(("num_triangles", "num_emitters", "num_bvh_nodes", "device_bytes"), (x.value for x in v)))

    def render(self, settings) -> dict:
        """settings: one abi.PtSettings per scene / view (total_spp iterations each)."""
        arr = (abi.PtSettings * len(self))(*settings)
        stats = abi.RayStats()
        self.device.check(self.device.lib.sspt_render_batch(self.device.ptr, self.ptr, arr, C.byref(stats)))
        return stats.as_dict()

    def render_adjoint(self, settings, dloss_dimages, first_iteration: int = 0) -> dict:
        """Adds dLoss/dtexel of every scene / view to the world's texture gradients.
        dloss_dimages: one (H, W, 3) array per scene / view (or all of them concatenated in
        one contiguous float32 torch tensor on the GPU); scene i replays
        iterations first_iteration .. first_iteration + settings[i].total_spp - 1."""
        arr = (abi.PtSettings * len(self))(*settings)
        expect = sum(h * w * 3 for h, w in self.sizes)
        if hasattr(dloss_dimages, "data_ptr") and getattr(dloss_dimages, "is_cuda", False):
            ptr, keep = _image_pointer(dloss_dimages, expect)  # all scenes' images, concatenated, on the GPU
        else:
            flat = np.concatenate([np.ascontiguousarray(g, np.float32).reshape(-1) for g in dloss_dimages])
            ptr, keep = _image_pointer(flat, expect)
        stats = abi.RayStats()
        self.device.check(self.device.lib.sspt_render_adjoint_batch(self.device.ptr, self.ptr, arr, ptr,
                                                                    first_iteration, C.byref(stats)))
        del keep
        return stats.as_dict()

    def read_layer(self, index: int, name: str) -> np.ndarray:
        h, w = self.sizes[index]
        return _read_layer(self.device, self.ptr, index, name, h, w)

    def track_outliers(self, n: int):
        """Outlier tracking of n <= 4 samples per pixel for every scene / view (0 = off)."""
        self.device.check(self.device.lib.sspt_framebuffer_track_outliers(self.device.ptr, self.ptr, n))
        self.num_outliers = n

    def read_outliers(self, index: int):
        h, w = self.sizes[index]
        return _read_outliers(self.device, self.ptr, index, h, w, self.num_outliers)

    def replay_pixel(self, index: int, settings: abi.PtSettings, col: int, row: int, iteration: int):
        """ReplayPixel of scene / view `index` of the world."""
        return _replay(self.device, self.ptr, index, settings, col, row, iteration)

    def read_image(self, index: int) -> np.ndarray:
        h, w = self.sizes[index]
        img = np.zeros((h, w, 3), np.float32)
        self.device.check(self.device.lib.sspt_framebuffer_read_scene(self.device.ptr, self.ptr, index,
                                                                      abi.fptr(img)))
        return img


_LAYERS = {"albedo": abi.LAYER_ALBEDO, "normal": abi.LAYER_NORMAL}


def _read_layer(device, ptr, index, name, h, w):
    if name not in _LAYERS:
        raise KeyError(f"unknown layer {name!r} (albedo, normal)")
    img = np.zeros((h, w, 3), np.float32)
    device.check(device.lib.sspt_framebuffer_read_layer(device.ptr, ptr, index, _LAYERS[name], abi.fptr(img)))
    return img


def _replay(device, ptr, index, settings, col, row, iteration):
    """sspt_replay_pixel: (node array, RGB estimate) of one path."""
    cap = 2 + settings.max_depth * (1 + 2 * max(0, settings.num_shadow_rays))
    nodes = (abi.PathNode * cap)()
    count = C.c_uint32()
    est = np.zeros(3, np.float32)
    device.check(device.lib.sspt_replay_pixel(device.ptr, ptr, index, C.byref(settings), col, row, iteration, nodes,
                                              cap, C.byref(count), abi.fptr(est)))
    if count.value > cap:
        raise RuntimeError(f"replayed path has {count.value} nodes, more than the bound {cap}")
    return list(nodes[:count.value]), est


def _read_outliers(device, ptr, index, h, w, n):
    weights = np.zeros((h, w, n), np.float32)
    its = np.zeros((h, w, n), np.int32)
    device.check(device.lib.sspt_framebuffer_read_outliers(device.ptr, ptr, index, abi.fptr(weights),
                                                           its.ctypes.data_as(C.POINTER(C.c_int32))))
    return weights, its


def _hits(hits):
    dt = np.dtype([("t", "<f4"), ("u", "<f4"), ("v", "<f4"), ("prim", "<i4")])
    return np.frombuffer(hits, dtype=dt).copy()


class Integrator:
    """SeeSharp/Integrators/Integrator.cs:6-30."""

    def __init__(self):
        self.MaxDepth = 100
        self.MinDepth = 1

    def Render(self, scene):  # pragma: no cover - abstract
        raise NotImplementedError

    def ReplayPixel(self, scene, pixel, iteration):
        raise NotImplementedError("Path replay is not supported by this integrator")


class GpuPathTracer(Integrator):
    """The unidirectional path tracer with NEE, balance-heuristic MIS and
    Russian roulette (SeeSharp/Integrators/PathTracer.cs) on one GPU."""

    def __init__(self, device: Device | None = None):
        super().__init__()
        self.BaseSeed = 0xC030114
        self.TotalSpp = 20
        self.MaximumRenderTimeMs = None
        self.NumShadowRays = 1
        self.EnableBsdfDI = True
        # EnableDenoiser (default true, PathTracer.cs:49) writes the "albedo" and
        # "normal" feature layers; running the OIDN denoiser on them is a separate
        # product and not part of this tracer. The technique pyramid is not supported.
        self.RenderTechniquePyramid = False
        self.EnableDenoiser = True
        self.device = device
        self._gpu_scene = None

    def settings(self) -> abi.PtSettings:
        return abi.PtSettings(
            self.BaseSeed & 0xFFFFFFFF, self.TotalSpp, self.MaxDepth, self.MinDepth, self.NumShadowRays,
            1 if self.EnableBsdfDI else 0,
            -1 if self.MaximumRenderTimeMs is None else int(self.MaximumRenderTimeMs),
            1 if self.EnableDenoiser else 0)

    def prepare(self, scene) -> GpuScene:
        """Scene.Prepare on the device (BVH build, emitter maps, env grid)."""
        if self.device is None:
            self.device = Device(0)
        if self._gpu_scene is None or self._gpu_scene.scene is not scene:
            self._gpu_scene = GpuScene(self.device, scene)
        return self._gpu_scene

    def Render(self, scene):
        if self.RenderTechniquePyramid:
            raise NotImplementedError("RenderTechniquePyramid is not supported by the GPU path tracer")
        gs = self.prepare(scene)
        fb = scene.FrameBuffer
        n_outl = max(0, min(abi.MAX_OUTLIERS, int(getattr(fb, "NumOutliersToTrack", 0))))
        if gs.num_outliers != n_outl:
            gs.track_outliers(n_outl)
        gs.clear()
        cfg = self.settings()
        if getattr(fb, "ReferenceImage", None) is None:
            # the library checks MaximumRenderTimeMs before every iteration
            t0 = time.perf_counter()
            stats = gs.render(cfg, 0, self.TotalSpp)
            render_ms = (time.perf_counter() - t0) * 1000.0
            iterations = stats["num_iterations"]
        else:
            # FrameBuffer.EndIteration with a reference image (FrameBuffer.cs:267-291):
            # one error entry per iteration, timed over the render calls only; the time
            # budget follows RenderTimer (PathTracer.cs:221-227, RenderTimer.cs:35-60):
            # whole milliseconds per iteration, and the next iteration only while the
            # time so far plus its average per iteration stays within the budget
            stats = {k: 0 for k in ("num_rays", "num_shadow_rays", "num_occluded", "num_ray_hits",
                                    "num_shadow_rays_traced")}
            render_ms = 0.0
            whole_ms = 0
            iterations = 0
            fb.Errors = []
            one = self.settings()
            one.max_render_time_ms = -1
            for it in range(self.TotalSpp):
                per_iteration = whole_ms // iterations if iterations else 0
                if self.MaximumRenderTimeMs is not None and whole_ms + per_iteration > self.MaximumRenderTimeMs:
                    break
                t0 = time.perf_counter()
                st = gs.render(one, it, 1)
                elapsed = (time.perf_counter() - t0) * 1000.0
                render_ms += elapsed
                whole_ms += int(elapsed)
                iterations += 1
                for k in stats:
                    stats[k] += st[k]
                fb.Errors.append(metrics.ErrorMetric(render_ms, gs.read_image(), fb.ReferenceImage))
        fb.Image = gs.read_image()
        if self.EnableDenoiser:
            fb.Layers["albedo"] = gs.read_layer("albedo")
            fb.Layers["normal"] = gs.read_layer("normal")
        fb.Outliers = gs.read_outliers() if n_outl else None
        fb.CurIteration = iterations
        fb.NaNWarnings = gs.nan_count()
        fb.MetaData["RenderTime"] = render_ms
        fb.MetaData["NumIterations"] = iterations
        fb.MetaData["RayTracerStats"] = {
            "NumRays": stats["num_rays"], "NumShadowRays": stats["num_shadow_rays"],
            "NumOccluded": stats["num_occluded"], "NumRayHits": stats["num_ray_hits"]}
        return fb

    def ReplayPixel(self, scene, pixel, iteration: int):
        """PathTracerBase.ReplayPixel (PathTracer.cs:61-69): re-traces the path of
        pixel (col, row) in `iteration` and returns (PathGraph, estimate); the
        framebuffer is not touched."""
        from .pathgraph import PathGraph

        col, row = pixel
        gs = self.prepare(scene)
        nodes, est = gs.replay_pixel(self.settings(), int(col), int(row), int(iteration))
        return PathGraph.from_nodes(nodes), est


class DebugVisualizer(Integrator):
    """A fast grayscale view of a scene (SeeSharp/Integrators/DebugVisualizer.cs:6-59):
    per pixel and sample the jittered primary hit shaded by |cos| between its
    face normal and the direction back to the camera ("eye light")."""

    def __init__(self, device: Device | None = None):
        super().__init__()
        self.BaseSeed = 0xC030114
        self.TotalSpp = 1
        self.device = device
        self._gpu_scene = None

    def Render(self, scene):
        if self.device is None:
            self.device = Device(0)
        if self._gpu_scene is None or self._gpu_scene.scene is not scene:
            self._gpu_scene = GpuScene(self.device, scene)
        gs = self._gpu_scene
        fb = scene.FrameBuffer
        gs.clear()
        t0 = time.perf_counter()
        gs.render_debug(self.BaseSeed, 0, self.TotalSpp)
        fb.MetaData["RenderTime"] = (time.perf_counter() - t0) * 1000.0
        fb.MetaData["NumIterations"] = self.TotalSpp
        fb.Image = gs.read_image()
        fb.CurIteration = self.TotalSpp
        return fb


class VertexCacheBidirLightTracer(Integrator):
    """Light tracing: a VertexCacheBidir (SeeSharp/Integrators/Bidir/VertexCacheBidir.cs)
    with NumConnections = 0, NumShadowRays = 0, EnableHitting = false and
    EnableLightTracer = true, rendered by BidirBase.Render
    (SeeSharp/Integrators/Bidir/BidirBase.cs:84-146). Every iteration traces
    NumLightPaths paths from the emitters and splats each vertex onto the camera
    (BidirBase.Light.cs:35-277). With a background, light paths start from it too and
    camera paths add its radiance where they leave the scene (BidirBase.OnBackgroundHit)."""

    def __init__(self, device: Device | None = None):
        super().__init__()
        self.NumIterations = 2
        self.NumLightPaths = -1
        self.BaseSeedLight = 0x13C0FEFE
        self.device = device
        self._gpu_scene = None

    def settings(self) -> abi.LtSettings:
        return abi.LtSettings(self.BaseSeedLight & 0xFFFFFFFF, int(self.NumLightPaths), self.MaxDepth,
                              self.MinDepth)

    def Render(self, scene):
        if self.device is None:
            self.device = Device(0)
        if self._gpu_scene is None or self._gpu_scene.scene is not scene:
            self._gpu_scene = GpuScene(self.device, scene)
        gs = self._gpu_scene
        fb = scene.FrameBuffer
        gs.clear()
        t0 = time.perf_counter()
        stats = gs.render_light_tracer(self.settings(), 0, self.NumIterations)
        fb.MetaData["RenderTime"] = (time.perf_counter() - t0) * 1000.0
        fb.MetaData["NumIterations"] = self.NumIterations
        fb.MetaData["RayTracerStats"] = {
            "NumRays": stats["num_rays"], "NumShadowRays": stats["num_shadow_rays"],
            "NumOccluded": stats["num_occluded"], "NumRayHits": stats["num_ray_hits"]}
        fb.Image = gs.read_image()
        fb.CurIteration = self.NumIterations
        fb.NaNWarnings = gs.nan_count()
        return fb


class ClassicBidir(Integrator):
    """Classic bidirectional path tracing (SeeSharp/Integrators/Bidir/ClassicBidir.cs)
    rendered by BidirBase.Render (SeeSharp/Integrators/Bidir/BidirBase.cs:84-146):
    every iteration traces one light path per pixel, splats its vertices onto
    the camera (EnableLightTracer), and traces one camera path per pixel that
    adds emitter hits, NumShadowRays next-event samples and, with
    EnableConnections, connections to every vertex of its pixel's light path,
    all combined by the balance heuristic. A background is a light of its own:
    light paths start from it (EnvironmentMap.SampleRay) with probability
    1 / (1 + emitters), next events sample it, camera paths that leave the
    scene hit it."""

    def __init__(self, device: Device | None = None):
        super().__init__()
        self.NumIterations = 2
        self.NumLightPaths = -1
        self.BaseSeedCamera = 0xC030114
        self.BaseSeedLight = 0x13C0FEFE
        self.NumShadowRays = 1
        self.EnableConnections = True
        self.EnableLightTracer = True
        self.device = device
        self._gpu_scene = None

    def settings(self) -> abi.BidirSettings:
        return abi.BidirSettings(self.BaseSeedCamera & 0xFFFFFFFF, self.BaseSeedLight & 0xFFFFFFFF, self.MaxDepth,
                                 self.MinDepth, self.NumShadowRays, 1 if self.EnableConnections else 0,
                                 1 if self.EnableLightTracer else 0, int(self.NumLightPaths))

    def Render(self, scene):
        if self.device is None:
            self.device = Device(0)
        if self._gpu_scene is None or self._gpu_scene.scene is not scene:
            self._gpu_scene = GpuScene(self.device, scene)
        gs = self._gpu_scene
        fb = scene.FrameBuffer
        gs.clear()
        t0 = time.perf_counter()
        stats = gs.render_bidir(self.settings(), 0, self.NumIterations)
        fb.MetaData["RenderTime"] = (time.perf_counter() - t0) * 1000.0
        fb.MetaData["NumIterations"] = self.NumIterations
        fb.MetaData["RayTracerStats"] = {
            "NumRays": stats["num_rays"], "NumShadowRays": stats["num_shadow_rays"],
            "NumOccluded": stats["num_occluded"], "NumRayHits": stats["num_ray_hits"]}
        fb.Image = gs.read_image()
        fb.CurIteration = self.NumIterations
        fb.NaNWarnings = gs.nan_count()
        return fb


class VertexCacheBidir(Integrator):
    """Bidirectional path tracing with a light vertex cache
    (SeeSharp/Integrators/Bidir/VertexCacheBidir.cs) rendered by BidirBase.Render:
    NumLightPaths light paths per iteration, their vertices splatted onto the
    camera (EnableLightTracer), and one camera path per pixel with emitter hits
    (EnableHitting), NumConnections connections to vertices drawn from the whole
    cache and NumShadowRays next-event samples, combined by the balance
    heuristic. Backgrounds as for ClassicBidir."""

    algorithm = abi.VERTEX_CACHE_BIDIR

    def __init__(self, device: Device | None = None):
        super().__init__()
        self.NumIterations = 2
        self.NumLightPaths = -1
        self.BaseSeedCamera = 0xC030114
        self.BaseSeedLight = 0x13C0FEFE
        self.NumConnections = 1
        self.NumShadowRays = 1
        self.EnableLightTracer = True
        self.EnableHitting = True
        self.device = device
        self._gpu_scene = None

    def settings(self) -> abi.VcmSettings:
        return abi.VcmSettings(self.algorithm, self.BaseSeedCamera & 0xFFFFFFFF, self.BaseSeedLight & 0xFFFFFFFF,
                               self.MaxDepth, self.MinDepth, int(self.NumLightPaths), self.NumConnections,
                               self.NumShadowRays, int(self.EnableLightTracer), int(self.EnableHitting),
                               int(getattr(self, "EnableMerging", False)), int(getattr(self, "MergePrimary", False)),
                               int(getattr(self, "MaxNumPhotons", 8)),
                               int(getattr(self, "DisableCorrelAwareMIS", False)))

    def Render(self, scene):
        if self.device is None:
            self.device = Device(0)
        if self._gpu_scene is None or self._gpu_scene.scene is not scene:
            self._gpu_scene = GpuScene(self.device, scene)
        gs = self._gpu_scene
        fb = scene.FrameBuffer
        gs.clear()
        t0 = time.perf_counter()
        stats = gs.render_vcm(self.settings(), 0, self.NumIterations)
        fb.MetaData["RenderTime"] = (time.perf_counter() - t0) * 1000.0
        fb.MetaData["NumIterations"] = self.NumIterations
        fb.MetaData["RayTracerStats"] = {
            "NumRays": stats["num_rays"], "NumShadowRays": stats["num_shadow_rays"],
            "NumOccluded": stats["num_occluded"], "NumRayHits": stats["num_ray_hits"]}
        fb.Image = gs.read_image()
        fb.CurIteration = self.NumIterations
        fb.NaNWarnings = gs.nan_count()
        return fb


class VertexConnectionAndMerging(VertexCacheBidir):
    """Vertex connection and merging (SeeSharp/Integrators/Bidir/VertexConnectionAndMerging.cs):
    VertexCacheBidir plus merging with the MaxNumPhotons nearest light vertices
    within the camera path's pixel footprint (EnableMerging, MergePrimary),
    weighted by the correlation-aware balance heuristic unless
    DisableCorrelAwareMIS."""

    algorithm = abi.VCM

    def __init__(self, device: Device | None = None):
        super().__init__(device)
        self.EnableMerging = True
        self.MergePrimary = False
        self.MaxNumPhotons = 8
        self.DisableCorrelAwareMIS = False
