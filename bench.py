"""Benchmark: Mpaths/s of the GPU path tracer on the 1M-triangle 1024^2 depth-8 scene.

One step = one full render of the large scene (1 036 296 triangles, 1024 x 1024,
64 spp, MaxDepth 8 = 67 108 864 paths) into the device framebuffer, with the
scene already resident in HBM (BVH build and upload happen before timing).

Prints one JSON line with the metric, the roofline of the dominant kernel
(from HIP events on the library's own stream), the metric's second half (the
adjoint / forward cost ratio on the 512^2 textured box and on 32 views of the
textured large scene), the batching form's speed-up (64 scenes in one world vs
per-scene launches) and the CPU baseline (the oracle restatement of the
reference, timed on a bounded pixel sample).

    python bench.py [--gpus 1] [--steps 3] [--warmup 1]
    python bench.py --workload batch     # 64 scenes in shared waves vs per-scene launches
    python bench.py --workload adjoint   # path-replay adjoint vs forward cost (512^2 textured box)
    python bench.py --workload views     # 32 views of the 1M-tri scene with 1024^2 ground textures: forward + adjoint
    python bench.py --workload copy      # plain 1 GiB device copy: HBM counter calibration for rocprofv3 --pmc
    python bench.py --workload cornell   # the reference's CornellBox.json 512^2 512 spp render vs its recorded times
    python bench.py --workload bvh       # device BVH build of the 1M-tri scene vs the host builder
    python bench.py --workload bidir     # ClassicBidir and the light tracer on the 1M-tri scene without its sky
"""
from __future__ import annotations

import argparse
import json
import os
import re
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = ("Mpaths/s at 1M-tri 1024² depth-8 PT (relMSE≤1e-3); adjoint/forward cost ratio")
HBM_PEAK_GBS = 8000.0

# Algorithmic bytes per unit (per ray / hit / path), split by the kernel that
# moves them: extension ray 32 B ray record + 16 B hit record (extend) and
# 64 B path state (shade); shadow ray 32 B + 4 B visibility (shadow); hit
# 3x12 B positions + 3x12 B normals + 3x8 B uvs + 4 B material (shade);
# path state written by generate and read by accumulate with the pixel.
BYTES_EXTEND_PER_RAY = 48
BYTES_SHADOW_PER_RAY = 36
BYTES_SHADE_PER_RAY = 64
BYTES_SHADE_PER_HIT = 100
BYTES_GENERATE_PER_PATH = 124   # queue entry (36 B) + path state (88 B)
BYTES_ACCUMULATE_PER_PATH = 72  # path radiance (48 B) + pixel read and write (24 B)
# the fused primary launch: path state (88 B) + queue entry without the origin (20 B) + hit record (16 B)
BYTES_PRIMARY_PER_PATH = 124
TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "pt_large_scene_hbm_traffic.md")
SQ_PROFILE = os.path.join(ROOT, "profiles", "pt_large_scene_sq_counters.md")


def measured_traffic():
    """HBM bytes per launch of each kernel on the default large-scene workload,
    from the committed rocprofv3 PMC summary (FETCH_SIZE / WRITE_SIZE scaled by
    the copy-kernel calibration). Empty if the summary is absent."""
    try:
        text = open(TRAFFIC_PROFILE).read()
    except OSError:
        return {}
    m = re.search(r"^traffic_bytes_per_launch:(.*)$", text, re.M)
    if not m:
        return {}
    return {k: int(v) for k, v in re.findall(r"(\w+)=(\d+)", m.group(1))}


def measured_share(key):
    """Per-kernel shares on the default large-scene workload from the committed
    rocprofv3 counter summary: "valu_issue_share" (VALU issue slots filled) or
    "td_busy_share" (the texture data units' busy cycles)."""
    try:
        text = open(SQ_PROFILE).read()
    except OSError:
        return {}
    m = re.search(r"^" + key + r":(.*)$", text, re.M)
    if not m:
        return {}
    return {k: float(v) for k, v in re.findall(r"(\w+)=([0-9.]+)", m.group(1))}


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def cpu_baseline(scene, spp, max_depth, budget_s):
    """Times the oracle (CPU restatement) on a seeded pixel subset of the same workload."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as ol

    # one GPU's share of the host: at most 16 threads whatever the host has
    threads = min(16, len(os.sched_getaffinity(0)))
    t0 = time.perf_counter()
    o = ol.OracleScene(scene)
    log(f"[bench] oracle scene ready in {time.perf_counter() - t0:.1f} s")
    npx = scene.FrameBuffer.Width * scene.FrameBuffer.Height
    rng = np.random.default_rng(12345)
    cfg = ol.settings(total_spp=spp, max_depth=max_depth)
    # calibrate on a small sample, then size the timed sample to the budget
    pix = np.sort(rng.choice(npx, 1024, replace=False)).astype(np.int32)
    t0 = time.perf_counter()
    o.render(cfg, 0, spp, pixels=pix, threads=threads)
    rate = 1024 * spp / (time.perf_counter() - t0)
    n = int(min(npx, max(1024, rate * budget_s / spp)))
    pix = np.sort(rng.choice(npx, n, replace=False)).astype(np.int32)
    t0 = time.perf_counter()
    o.render(cfg, 0, spp, pixels=pix, threads=threads)
    dt = time.perf_counter() - t0
    o.close()
    value = n * spp / dt / 1e6
    host = host_cpu()
    return {"value": value, "unit": "Mpaths/s", "cores": threads, "kind": "port",
            "sample": f"{n} seeded pixels x {spp} spp of the same scene ({n * spp} paths, {dt:.1f} s)",
            "per_core_mpaths_per_s": round(value / threads, 4),
            "host_cpu_model": host["model"], "host_logical_cpus": host["logical"],
            "host_cpus_available": host["available"],
            "note": "threads = one GPU's share of the host (16 at most); per_core x host_logical_cpus "
                    "bounds what the whole host would reach"}


def host_cpu():
    """CPU model (lscpu's "Model name", from /proc/cpuinfo) and logical CPU counts."""
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.lower().startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"model": model, "logical": os.cpu_count(), "available": len(os.sched_getaffinity(0))}


def measure_batch(dev, steps, autotune_first=False):
    """The batching workload: 64 scenes (seeds 1000..1063, 1.8k-194k triangles,
    128^2-512^2, 4/16/64 spp, MaxDepth 8) rendered in shared waves, against the
    same 64 scenes rendered one sspt_render call after the other, with the
    wavefront and with the plain single-kernel tracer."""
    from seesharp_amd import GpuBatch, GpuScene, abi, autotune, scenes

    t0 = time.perf_counter()
    sc, spp = scenes.batch_scenes(64, 1000)
    log(f"[bench] 64 scenes generated in {time.perf_counter() - t0:.1f} s")
    cfgs = [abi.PtSettings(0xC030114, n, 8, 1, 1, 1, -1) for n in spp]
    npaths = sum(s.FrameBuffer.Width * s.FrameBuffer.Height * n for s, n in zip(sc, spp))
    t0 = time.perf_counter()
    batch = GpuBatch(dev, sc)
    info = batch.info()
    log(f"[bench] batch world ready in {time.perf_counter() - t0:.1f} s: {info}")
    tuned = None
    if autotune_first:
        tuned, _ = autotune.tune(dev, lambda: batch.render(cfgs), capacities=(1 << 22, 1 << 23, 1 << 24), log=log)
        log(f"[bench] autotuned: {tuned}")
        tuned = {k: tuned[k] for k in ("trace_block", "shade_block", "wave_capacity", "trace_blocks", "shade_blocks",
                                       "ms")}
    batch.render(cfgs)  # warm-up
    t0 = time.perf_counter()
    for _ in range(steps):
        batch.render(cfgs)
    batch_ms = (time.perf_counter() - t0) * 1000.0 / steps
    log(f"[bench] batched: {batch_ms:.1f} ms per pass")
    # per-scene launches with the same execution settings
    singles = [GpuScene(dev, s) for s in sc]

    def per_scene():
        for g, c, n in zip(singles, cfgs, spp):
            g.clear()
            g.render(c, 0, n)

    per_scene()
    t0 = time.perf_counter()
    for _ in range(steps):
        per_scene()
    single_ms = (time.perf_counter() - t0) * 1000.0 / steps
    log(f"[bench] per-scene launches: {single_ms:.1f} ms per pass")
    # per-scene plain launches: the single-kernel tracer, one render per scene
    dev.set_execution_mode(True)
    try:
        per_scene()
        t0 = time.perf_counter()
        per_scene()
        plain_ms = (time.perf_counter() - t0) * 1000.0
    finally:
        dev.set_execution_mode(False)
    log(f"[bench] per-scene plain single-kernel launches: {plain_ms:.1f} ms per pass")
    for g in singles:
        g.close()
    batch.close()
    return {
        "workload": "64 scenes, seeds 1000..1063, triangle budget log-uniform [1k, 200k], 128^2..512^2, "
                    "spp in {4,16,64}, MaxDepth 8",
        "triangles": info["num_triangles"], "paths_per_pass": npaths,
        "batched_ms": round(batch_ms, 3),
        "batched_mpaths_per_s": round(npaths / (batch_ms / 1000.0) / 1e6, 3),
        "per_scene_launches_mpaths_per_s": round(npaths / (single_ms / 1000.0) / 1e6, 3),
        "speedup_vs_per_scene": round(single_ms / batch_ms, 3),
        "per_scene_plain_single_kernel_mpaths_per_s": round(npaths / (plain_ms / 1000.0) / 1e6, 3),
        "speedup_vs_per_scene_plain": round(plain_ms / batch_ms, 3),
        "autotuned": tuned,
    }


def batch_main(args):
    from seesharp_amd import Device

    r = measure_batch(Device(0), args.steps, not args.no_autotune)
    out = {
        "metric": "Mpaths/s of the batched wavefront (64 scenes, one world) vs per-scene launches",
        "value": r["batched_mpaths_per_s"], "unit": "Mpaths/s", "n_gpus": 1, "steps": args.steps, "warmup": 1,
        "ms_per_step": r["batched_ms"], "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic",
        "config": {"workload": r["workload"], "triangles": r["triangles"], "paths_per_step": r["paths_per_pass"]},
    }
    out.update({k: v for k, v in r.items() if k not in ("workload", "triangles", "paths_per_pass")})
    print(json.dumps(out), flush=True)


def measure_adjoint(dev, steps, kernels=True):
    """The differentiation workload: the box at 512^2 with 256^2 RGB albedo and
    mono roughness textures on the back wall and floor, forward 32 spp and
    path-replay adjoint 32 spp of an L2 loss, MaxDepth 6. The adjoint / forward
    cost ratio compares the wall time of one render call each (graphs on)."""
    from seesharp_amd import GpuScene, abi, scenes

    res, tex_res, spp, depth = 512, 256, 32, 6
    target_scene = scenes.box_scene(res, res, textured=True, tex_res=tex_res)
    tg = GpuScene(dev, target_scene)
    cfg = abi.PtSettings(0xC030114, spp, depth, 1, 1, 1, -1)
    tg.clear()
    tg.render(cfg, 0, 128)
    target = tg.read_image()
    tg.close()
    s = scenes.box_scene(res, res, textured=True, tex_res=tex_res, uniform_textures=0.5)
    g = GpuScene(dev, s)
    npaths = res * res * spp

    def forward():
        g.clear()
        return g.render(cfg, 0, spp)

    def adjoint(dimg):
        g.clear_gradients()
        return g.render_adjoint(cfg, dimg, spp, spp)

    forward()
    # dLoss/dImage of an L2 loss, resident on the GPU as in an optimisation loop (the library
    # copies it device to device)
    import torch

    dimg = torch.from_numpy((2.0 * (g.read_image() - target) / target.size).astype(np.float32)).to("cuda:0")
    adjoint(dimg)
    fwd_ms, adj_ms = [], []
    for _ in range(max(1, steps)):
        t0 = time.perf_counter()
        forward()
        fwd_ms.append((time.perf_counter() - t0) * 1000.0)
        t0 = time.perf_counter()
        stats = adjoint(dimg)
        adj_ms.append((time.perf_counter() - t0) * 1000.0)
    f_ms, a_ms = float(np.median(fwd_ms)), float(np.median(adj_ms))
    log(f"[bench] textured box: forward {f_ms:.2f} ms, adjoint {a_ms:.2f} ms per {npaths} paths")
    out = {
        "workload": "box scene 512x512, 256x256 RGB albedo + mono roughness textures on back wall and floor, "
                    "forward 32 spp + adjoint 32 spp, MaxDepth 6",
        "paths_per_pass": npaths,
        "ratio": round(a_ms / f_ms, 3),
        "forward_ms": round(f_ms, 3),
        "adjoint_ms": round(a_ms, 3),
        "forward_mpaths_per_s": round(npaths / f_ms / 1e3, 3),
        "adjoint_mpaths_per_s": round(npaths / a_ms / 1e3, 3),
        "adjoint_ray_stats": stats,
    }
    if kernels:
        # per-kernel device time, one event-timed render of each
        dev.set_kernel_timing(True)
        dev.reset_kernel_timing()
        forward()
        out["kernel_ms_forward"] = {k: round(dev.kernel_timing(k)[1], 3)
                                    for k in ("extend", "trace", "shade", "shadow", "accumulate")}
        dev.reset_kernel_timing()
        adjoint(dimg)
        out["kernel_ms_adjoint"] = {k: round(dev.kernel_timing(k)[1], 3)
                                    for k in ("extend", "trace_adjoint", "shade_adjoint", "shade_grad",
                                              "shadow_adjoint", "adjoint_finalize")}
        dev.set_kernel_timing(False)
    g.close()
    return out


def adjoint_main(args):
    from seesharp_amd import Device

    r = measure_adjoint(Device(0), args.steps)
    out = {
        "metric": "adjoint/forward cost ratio (path-replay adjoint of an L2 image loss)",
        "value": r["ratio"], "unit": "x", "n_gpus": 1, "steps": args.steps, "warmup": 1,
        "ms_per_step": round(r["forward_ms"] + r["adjoint_ms"], 3), "higher_is_better": False, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": r["workload"], "paths_per_pass": r["paths_per_pass"]},
    }
    out.update({k: v for k, v in r.items() if k not in ("workload", "paths_per_pass", "ratio")})
    print(json.dumps(out), flush=True)


def measure_views(dev, steps):
    """Differentiation + batching: the large scene with 1024^2 RGB albedo and
    1024^2 roughness textures on the ground grid, 32 views on a circle (radius 8,
    height 2) at 512^2 and 16 spp, MaxDepth 8. All views go forward in one
    batched render and backward in one batched adjoint whose gradients add into
    the shared texture buffers; per-view L2 loss against targets rendered from
    seeded noise textures (seed 5). Initial textures are uniform 0.5."""
    from seesharp_amd import GpuBatch, abi, scenes

    res, spp, depth, nviews = 512, 16, 8, 32
    t0 = time.perf_counter()
    target_scene = scenes.large_scene(res, res, ground_textures=1024, ground_texture_seed=5)
    s = scenes.large_scene(res, res, ground_textures=1024, ground_uniform=0.5)
    cams = scenes.circle_views(nviews, radius=8.0, height=2.0)
    views = GpuBatch.views(dev, s, cams)
    log(f"[bench] view set ready in {time.perf_counter() - t0:.1f} s: {views.info()}")
    cfgs = [abi.PtSettings(0xC030114, spp, depth, 1, 1, 1, -1) for _ in cams]
    p = target_scene.Meshes[-5].Material.MaterialParameters  # the ground (before the 4 light quads)
    q = s.Meshes[-5].Material.MaterialParameters
    idx_alb, idx_rough = views.texture_index(q.BaseColor), views.texture_index(q.Roughness)
    views.write_texture(idx_alb, p.BaseColor.Image)
    views.write_texture(idx_rough, p.Roughness.Image)
    views.render([abi.PtSettings(571298512, 64, depth, 1, 1, 1, -1) for _ in cams])
    targets = [views.read_image(i) for i in range(nviews)]
    views.write_texture(idx_alb, q.BaseColor.Image)
    views.write_texture(idx_rough, q.Roughness.Image)
    npaths = nviews * res * res * spp

    def adjoint(dl):
        views.clear_gradients()
        return views.render_adjoint(cfgs, dl, first_iteration=spp)

    views.render(cfgs)
    imgs = [views.read_image(i) for i in range(nviews)]
    npx = res * res * 3
    # dLoss/dImage of every view, resident on the GPU as in an optimisation loop
    import torch

    dl = torch.from_numpy(np.concatenate([(2.0 * (imgs[i] - targets[i]) / (npx * nviews)).astype(np.float32).ravel()
                                          for i in range(nviews)])).to("cuda:0")
    adjoint(dl)
    fwd_ms, adj_ms = [], []
    for _ in range(max(1, steps)):
        t0 = time.perf_counter()
        views.render(cfgs)
        fwd_ms.append((time.perf_counter() - t0) * 1000.0)
        t0 = time.perf_counter()
        stats = adjoint(dl)
        adj_ms.append((time.perf_counter() - t0) * 1000.0)
    f_ms, a_ms = float(np.median(fwd_ms)), float(np.median(adj_ms))
    g_alb = views.gradient(idx_alb)
    log(f"[bench] 32 views: forward {f_ms:.1f} ms, adjoint {a_ms:.1f} ms per {npaths} paths")
    views.close()
    return {
        "workload": "large scene (1036296 tris) with 1024x1024 RGB albedo + mono roughness on the ground, "
                    "32 views on a circle (r 8, h 2) at 512x512, 16 spp, MaxDepth 8, L2 loss",
        "paths_per_pass": npaths,
        "ratio": round(a_ms / f_ms, 3),
        "forward_ms": round(f_ms, 3),
        "adjoint_ms": round(a_ms, 3),
        "forward_mpaths_per_s": round(npaths / f_ms / 1e3, 3),
        "adjoint_mpaths_per_s": round(npaths / a_ms / 1e3, 3),
        "albedo_gradient_abs_sum": float(np.abs(g_alb).sum()),
        "adjoint_ray_stats": stats,
    }


def views_main(args):
    from seesharp_amd import Device

    r = measure_views(Device(0), args.steps)
    out = {
        "metric": "adjoint/forward cost ratio, 32 views in one batched launch (path-replay adjoint)",
        "value": r["ratio"], "unit": "x", "n_gpus": 1, "steps": args.steps, "warmup": 1,
        "ms_per_step": round(r["forward_ms"] + r["adjoint_ms"], 3), "higher_is_better": False, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": r["workload"], "paths_per_pass": r["paths_per_pass"]},
    }
    out.update({k: v for k, v in r.items() if k not in ("workload", "paths_per_pass", "ratio")})
    print(json.dumps(out), flush=True)


def copy_main(args):
    """A plain elementwise device copy of known size (reads 1 GiB, writes 1 GiB
    per launch). Under rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE it calibrates
    what the counters report per byte moved, so that the path tracer's kernel
    counters can be converted to HBM bytes."""
    import torch

    n = 1 << 28  # float32 elements = 1 GiB
    src = torch.ones(n, dtype=torch.float32, device="cuda:0")
    dst = torch.empty_like(src)
    torch.mul(src, 1.0, out=dst)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(max(1, args.steps)):
        torch.mul(src, 1.0, out=dst)
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1000.0 / max(1, args.steps)
    print(json.dumps({"metric": "device copy bandwidth (read + write)", "value": round(2 * n * 4 / ms / 1e6, 1),
                      "unit": "GB/s", "bytes_read_per_launch": n * 4, "bytes_written_per_launch": n * 4,
                      "ms_per_launch": round(ms, 3), "launches": max(1, args.steps) + 1}), flush=True)


CORNELL_SCENE = os.path.join(ROOT, "tests", "golden", "scenes", "CornellBox", "CornellBox.json")
# Render times recorded with the reference's CornellBox reference renders
# (Data/Scenes/CornellBox/References/MaxDepth{5,2}-Width512-Height512.json:2),
# configured by Data/Scenes/CornellBox/References/Config.json:1-8.
CORNELL_REFERENCE_MS = {5: 142532.0, 2: 26768.0}


def bvh_main(args):
    """The scene build's BVH step (CommitScene, SeeSharp/Scene.cs:108-113) on
    the 1 036 296-triangle large scene: the device builder timed `steps` times
    after `warmup` builds (summed kernel time from HIP events, and wall time of
    the whole build incl. the triangle upload and node download), the host
    builder once, and whether the two trees are identical."""
    from seesharp_amd import Device, GpuScene, scenes

    dev = Device(0)
    scene = scenes.large_scene(args.res, args.res)
    for i in range(max(0, args.warmup)):
        GpuScene(dev, scene).close()
    kernel, wall = [], []
    g = None
    for i in range(max(1, args.steps)):
        if g is not None:
            g.close()
        g = GpuScene(dev, scene)
        t = g.bvh_build_time()
        kernel.append(t["kernel_ms"])
        wall.append(t["wall_ms"])
        log(f"[bench] device BVH build {i + 1}: kernels {t['kernel_ms']:.2f} ms, wall {t['wall_ms']:.1f} ms")
    dev.set_bvh_builder(host=True)
    h = GpuScene(dev, scene)
    dev.set_bvh_builder(host=False)
    host_wall = h.bvh_build_time()["wall_ms"]
    nd, td = g.read_bvh()
    nh, th = h.read_bvh()
    same = bool(nd.shape == nh.shape and (nd[:, :24] == nh[:, :24]).all()
                and np.array_equal(nd[:, 24:].view(np.int32), nh[:, 24:].view(np.int32))
                and np.array_equal(td, th))
    info = g.info()
    out = {
        "metric": "BVH build time, 1M-tri large scene (device builder kernels)",
        "value": round(float(np.mean(kernel)), 3),
        "unit": "ms",
        "n_gpus": 1,
        "steps": max(1, args.steps),
        "warmup": args.warmup,
        "ms_per_step": round(float(np.mean(kernel)), 3),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"large scene ({info['num_triangles']} triangles), binned SAH 64 bins, "
                               f"leaves of at most {dev.bvh_leaf_size()} triangles, 4-wide collapse"},
        "bvh_nodes": info["num_bvh_nodes"],
        "device_wall_ms": round(float(np.mean(wall)), 2),
        "host_builder_wall_ms": round(host_wall, 2),
        "trees_identical": same,
    }
    print(json.dumps(out), flush=True)


def bidir_main(args):
    """Bidirectional rendering on the 1 036 296-triangle large scene without its
    sky (its area lights only), 1024^2, MaxDepth 8: ClassicBidir with every
    technique on (one light path and one camera path per pixel and iteration,
    NumShadowRays 1), `spp` iterations per step; the light tracer alone
    (VertexCacheBidir as a pure light tracer), VertexCacheBidir and VCM (one
    connection, 8 photons per merge) beside it. Kernel times from HIP events; the
    oracle's ClassicBidir timed on one iteration of the same scene."""
    from seesharp_amd import Device, GpuScene, abi, scenes

    dev = Device(0)
    scene = scenes.large_scene(args.res, args.res, with_env=False)
    g = GpuScene(dev, scene)
    iters = max(1, args.spp // 8)
    cfg = abi.BidirSettings(0xC030114, 0x13C0FEFE, args.max_depth, 1, 1, 1, 1, -1)
    for _ in range(max(0, args.warmup)):
        g.clear()
        g.render_bidir(cfg, 0, iters)
    dev.set_kernel_timing(True)
    dev.reset_kernel_timing()
    ms = []
    for _ in range(max(1, args.steps)):
        g.clear()
        t0 = time.perf_counter()
        st = g.render_bidir(cfg, 0, iters)
        ms.append((time.perf_counter() - t0) * 1000.0)
    kt = {k: dev.kernel_timing(k) for k in ("bidir_light", "bidir_camera")}
    dev.reset_kernel_timing()
    lt = abi.LtSettings(0x13C0FEFE, -1, args.max_depth, 1)
    g.clear()
    g.render_light_tracer(lt, 0, 1)
    dev.reset_kernel_timing()
    g.clear()
    lst = g.render_light_tracer(lt, 0, iters)
    lt_ms = dev.kernel_timing("light_paths")
    # VertexCacheBidir and VCM (one connection, one shadow ray, 8 photons per merge) on the same scene
    npx = args.res * args.res
    vc = {}
    for name, algo in (("vertex_cache_bidir", abi.VERTEX_CACHE_BIDIR), ("vcm", abi.VCM)):
        vcfg = abi.VcmSettings(algo, 0xC030114, 0x13C0FEFE, args.max_depth, 1, -1, 1, 1, 1, 1, 1, 0, 8, 0)
        g.clear()
        g.render_vcm(vcfg, 0, 1)
        dev.reset_kernel_timing()
        g.clear()
        t0 = time.perf_counter()
        vst = g.render_vcm(vcfg, 0, iters)
        v_ms = (time.perf_counter() - t0) * 1000.0 / iters
        vc[name] = {"ms_per_iteration": round(v_ms, 3), "msamples_per_s": round(npx / v_ms / 1e3, 2),
                    "kernel_ms_per_iteration": {k: round(v[1] / max(1, v[0]), 3) for k, v in
                                                ((k, dev.kernel_timing(k)) for k in
                                                 ("vcm_light", "vcm_prepare", "vcm_camera"))},
                    "ray_stats_per_step": vst}
    dev.set_kernel_timing(False)
    npx = args.res * args.res
    step_ms = float(np.mean(ms))
    rays = st["num_rays"] + st["num_shadow_rays"]
    log(f"[bench] ClassicBidir {iters} iterations: {step_ms:.2f} ms; light kernel "
        f"{kt['bidir_light'][1] / max(1, kt['bidir_light'][0]):.2f} ms, camera kernel "
        f"{kt['bidir_camera'][1] / max(1, kt['bidir_camera'][0]):.2f} ms per iteration")
    out = {
        "metric": "ClassicBidir samples per second (one light + one camera path per pixel sample)",
        "value": round(npx * iters / step_ms / 1e3, 3),
        "unit": "Msamples/s",
        "n_gpus": 1,
        "steps": max(1, args.steps),
        "warmup": args.warmup,
        "ms_per_step": round(step_ms, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"large scene ({g.info()['num_triangles']} triangles) without its sky, "
                               f"{args.res}x{args.res}, ClassicBidir MaxDepth {args.max_depth}, connections + "
                               f"light tracer + 1 shadow ray, {iters} iterations per step",
                   "max_depth": args.max_depth, "iterations_per_step": iters},
        "ray_stats_per_step": st,
        "mrays_per_s": round(rays / step_ms / 1e3, 1),
        "kernel_ms_per_iteration": {k: round(v[1] / max(1, v[0]), 3) for k, v in kt.items()},
        "light_tracer": {"ms_per_iteration": round(lt_ms[1] / max(1, lt_ms[0]), 3),
                         "mpaths_per_s": round(npx * lt_ms[0] / lt_ms[1] / 1e3, 2) if lt_ms[1] else None,
                         "ray_stats_per_step": lst},
        **vc,
    }
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import oracle_lib as ol

        threads = min(16, len(os.sched_getaffinity(0)))
        o = ol.OracleScene(scene)
        t0 = time.perf_counter()
        o.render_bidir(ol.bidir_settings(max_depth=args.max_depth), 0, 1, threads=threads)
        dt = time.perf_counter() - t0
        o.close()
        out["cpu_baseline"] = {"value": round(npx / dt / 1e6, 4), "unit": "Msamples/s", "cores": threads,
                               "kind": "port", "sample": f"one iteration of the same render ({npx} samples, "
                                                         f"{dt:.1f} s)"}
    print(json.dumps(out), flush=True)


def cornell_main(args):
    """The reference's recorded render: CornellBox.json (36 triangles) at 512^2,
    512 spp, BaseSeed 571298512, MaxDepth 5 and 2, denoiser feature layers on
    (the reference's default). Timed: the render loop only, scene already on the
    device, as the reference's RenderTimer times it."""
    from seesharp_amd import Device, GpuScene, abi, sceneio

    dev = Device(0)
    scene = sceneio.load_scene(CORNELL_SCENE, 512, 512)
    g = GpuScene(dev, scene)
    spp = 512
    res = {}
    for depth in (5, 2):
        cfg = abi.PtSettings(571298512, spp, depth, 1, 1, 1, -1, 1)
        for _ in range(max(0, args.warmup)):
            g.clear()
            g.render(cfg, 0, spp)
        ms = []
        for _ in range(max(1, args.steps)):
            g.clear()
            t0 = time.perf_counter()
            st = g.render(cfg, 0, spp)
            ms.append((time.perf_counter() - t0) * 1000.0)
        res[depth] = (float(np.mean(ms)), st)
        log(f"[bench] CornellBox MaxDepth {depth}: {res[depth][0]:.2f} ms per {spp} spp")
    ms5, st5 = res[5]
    npaths = 512 * 512 * spp
    out = {
        "metric": "CornellBox 512x512 512spp MaxDepth 5 render time",
        "value": round(ms5, 3),
        "unit": "ms",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms5, 3),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "the reference's CornellBox.json scene file",
        "config": {"workload": "CornellBox.json (36 triangles, 2 emitters) 512x512, 512 spp, BaseSeed 571298512, "
                               "NEE + balance MIS, denoiser feature layers on",
                   "max_depth": 5, "paths_per_step": npaths},
        "mpaths_per_s": round(npaths / ms5 / 1e3, 3),
        "ray_stats_per_step": st5,
        "speedup_vs_reference_render_time": round(CORNELL_REFERENCE_MS[5] / ms5, 1),
        "max_depth_2": {"ms": round(res[2][0], 3), "reference_ms": CORNELL_REFERENCE_MS[2],
                        "speedup_vs_reference_render_time": round(CORNELL_REFERENCE_MS[2] / res[2][0], 1)},
    }
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--spp", type=int, default=64)
    ap.add_argument("--res", type=int, default=1024)
    ap.add_argument("--max-depth", type=int, default=8)
    ap.add_argument("--cpu-budget", type=float, default=15.0, help="seconds of CPU baseline work")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extras", action="store_true",
                    help="skip the adjoint (textured box, 32 views) and batching measurements")
    ap.add_argument("--workload", choices=("large", "batch", "adjoint", "views", "copy", "cornell", "bvh", "bidir"),
                    default="large")
    ap.add_argument("--no-autotune", action="store_true",
                    help="batching: keep the default block shape, wave capacity and grids instead of tuning them")
    args = ap.parse_args()
    # torch's HIP runtime must initialise before libsspt.so brings in the system one: the
    # adjoint measurements hand the library device tensors
    import torch  # noqa: F401

    if args.gpus != 1:
        print(f"bench.py runs on one GPU only; --gpus {args.gpus} is not supported "
              "(this build renders on a single GPU; sharding by iteration is a design, not code).", file=sys.stderr)
        sys.exit(2)

    if args.workload == "batch":
        batch_main(args)
        return
    if args.workload == "adjoint":
        adjoint_main(args)
        return
    if args.workload == "copy":
        copy_main(args)
        return
    if args.workload == "views":
        views_main(args)
        return
    if args.workload == "cornell":
        cornell_main(args)
        return
    if args.workload == "bidir":
        return bidir_main(args)
    if args.workload == "bvh":
        bvh_main(args)
        return

    from seesharp_amd import Device, GpuScene, scenes
    from seesharp_amd import abi

    t0 = time.perf_counter()
    scene = scenes.large_scene(args.res, args.res)
    log(f"[bench] scene generated: {scene.NumTriangles} triangles in {time.perf_counter() - t0:.1f} s")
    dev = Device(0)
    t0 = time.perf_counter()
    gs = GpuScene(dev, scene)
    info = gs.info()
    log(f"[bench] device scene ready in {time.perf_counter() - t0:.1f} s: {info}")
    cfg = abi.PtSettings(0xC030114, args.spp, args.max_depth, 1, 1, 1, -1)
    npaths = args.res * args.res * args.spp

    for i in range(args.warmup):
        gs.clear()
        gs.render(cfg, 0, args.spp)
        log(f"[bench] warmup {i + 1}/{args.warmup} done")
    # timed steps: waves replayed from a captured hipGraph, no per-kernel events
    totals = {"num_rays": 0, "num_shadow_rays": 0, "num_occluded": 0, "num_ray_hits": 0,
              "num_shadow_rays_traced": 0}
    step_ms = []
    for i in range(args.steps):
        gs.clear()
        t0 = time.perf_counter()
        st = gs.render(cfg, 0, args.spp)
        step_ms.append((time.perf_counter() - t0) * 1000.0)
        for k in totals:
            totals[k] += st[k]
        log(f"[bench] step {i + 1}/{args.steps}: {step_ms[-1]:.1f} ms")
    img = gs.read_image()
    ms = float(np.mean(step_ms))
    value = npaths / (ms / 1000.0) / 1e6
    # one more render with HIP events around every kernel launch (direct launches)
    dev.set_kernel_timing(True)
    dev.reset_kernel_timing()
    gs.clear()
    t0 = time.perf_counter()
    gs.render(cfg, 0, args.spp)
    timed_ms = (time.perf_counter() - t0) * 1000.0
    kernels = {k: dev.kernel_timing(k) for k in ("generate", "extend", "trace", "shade", "shadow", "accumulate")}
    kernels = {k: v for k, v in kernels.items() if v[0] > 0}
    dev.set_kernel_timing(False)
    log(f"[bench] event-timed render: {timed_ms:.1f} ms")

    n_ext = totals["num_rays"]
    n_sh = totals["num_shadow_rays_traced"]
    n_hit = totals["num_ray_hits"]
    # per render (the event-timed render is one render)
    per = 1.0 / args.steps
    alg = {
        "shade": (BYTES_SHADE_PER_RAY * n_ext + BYTES_SHADE_PER_HIT * n_hit) * per,
        "accumulate": BYTES_ACCUMULATE_PER_PATH * npaths,
        "generate": BYTES_GENERATE_PER_PATH * npaths,
    }
    if "trace" in kernels:
        # fused schedule: "extend" traces the primary rays, "trace" each later
        # bounce's continuation rays with the previous bounce's shadow rays, and
        # "shadow" the last bounce's shadow queue (empty: the last bounce samples
        # no next events)
        # "extend" starts the paths and traces their camera rays in one launch
        alg["extend"] = BYTES_PRIMARY_PER_PATH * npaths
        alg["trace"] = BYTES_EXTEND_PER_RAY * (n_ext * per - npaths) + BYTES_SHADOW_PER_RAY * n_sh * per
        alg["shadow"] = 0.0
    else:
        alg["extend"] = BYTES_EXTEND_PER_RAY * n_ext * per
        alg["shadow"] = BYTES_SHADOW_PER_RAY * n_sh * per
    dom = max(kernels, key=lambda k: kernels[k][1])
    launches, tot_ms = kernels[dom]
    avg_ms = tot_ms / max(1, launches)
    bytes_per_launch = alg[dom] / max(1, launches)
    achieved = bytes_per_launch / (avg_ms / 1000.0) / 1e9
    kernel_ms = sum(v[1] for v in kernels.values())
    default_workload = (args.res, args.spp, args.max_depth) == (1024, 64, 8)
    traffic = measured_traffic().get(dom) if default_workload else None
    valu_share = measured_share("valu_issue_share").get(dom) if default_workload else None
    td_share = measured_share("td_busy_share").get(dom) if default_workload else None

    out = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "Mpaths/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "large scene: 400 icospheres + 512x512-quad fBm ground + 4 quad lights + 512x256 sky, "
                        f"{info['num_triangles']} triangles, {args.res}x{args.res}, {args.spp} spp, "
                        f"MaxDepth {args.max_depth}, NEE + balance MIS, Russian roulette",
            "triangles": info["num_triangles"],
            "resolution": [args.res, args.res],
            "spp": args.spp,
            "max_depth": args.max_depth,
            "paths_per_step": npaths,
        },
        "mrays_per_s": round((n_ext + n_sh) / args.steps / (ms / 1000.0) / 1e6, 2),
        "ray_stats_per_step": {k: v // args.steps for k, v in totals.items()},
        "kernel_ms_per_step": {k: round(v[1], 3) for k, v in kernels.items()},
        "kernel_launches_per_step": {k: int(v[0]) for k, v in kernels.items()},
        "alg_bytes_per_step": {k: int(alg[k]) for k in kernels if k in alg},
        "event_timed_render_ms": round(timed_ms, 3),
        "kernel_sum_ms_per_step": round(kernel_ms, 3),
        "roofline": {
            "kernel": dom,
            "bound": "hbm",
            "achieved": round(achieved, 2),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 5),
            "traffic": traffic,
            "avg_launch_ms": round(avg_ms, 4),
            "launches": launches,
            "alg_bytes_per_launch": int(bytes_per_launch),
            # what does bound the kernel (rocprofv3 counters of this command; the HBM
            # fraction above is far from 1): its VALU instructions' share of the issue
            # slots and the texture data units' busy share (scattered node loads)
            "limiter": "valu_issue" if valu_share else None,
            "valu_issue_share": valu_share,
            "td_busy_share": td_share,
        },
        "image_mean": float(img.mean()),
        "nan_paths": gs.nan_count(),
        # scene preparation, outside the timed region: the BVH build (device builder)
        "scene_build": {"bvh_" + k: (round(v, 3) if isinstance(v, float) else v)
                        for k, v in gs.bvh_build_time().items()},
    }
    # the plain single-kernel tracer (one thread per path) on the same render
    dev.set_execution_mode(True)
    gs.clear()
    gs.render(cfg, 0, args.spp)  # warm-up / graph capture
    gs.clear()
    t0 = time.perf_counter()
    gs.render(cfg, 0, args.spp)
    single_ms = (time.perf_counter() - t0) * 1000.0
    dev.set_execution_mode(False)
    log(f"[bench] single-kernel tracer: {single_ms:.1f} ms")
    out["plain_single_kernel"] = {"mpaths_per_s": round(npaths / (single_ms / 1000.0) / 1e6, 3),
                                  "ms_per_step": round(single_ms, 3),
                                  "wavefront_speedup": round(single_ms / ms, 3)}
    if default_workload and not args.no_extras:
        # the metric's second half and the batching form, each timed in this run
        adj = measure_adjoint(dev, 3, kernels=False)
        vw = measure_views(dev, 1)
        bt = measure_batch(dev, 1, autotune_first=not args.no_autotune)
        out["adjoint_forward_ratio"] = {"textured_box_512": adj["ratio"], "views_32x512_large": vw["ratio"]}
        out["adjoint"] = {
            "textured_box_512": {k: adj[k] for k in ("workload", "forward_ms", "adjoint_ms", "paths_per_pass")},
            "views_32x512_large": {k: vw[k] for k in ("workload", "forward_ms", "adjoint_ms", "paths_per_pass")},
        }
        out["batching"] = {k: bt[k] for k in ("workload", "triangles", "paths_per_pass", "batched_mpaths_per_s",
                                              "per_scene_launches_mpaths_per_s", "speedup_vs_per_scene",
                                              "per_scene_plain_single_kernel_mpaths_per_s",
                                              "speedup_vs_per_scene_plain", "autotuned")}
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(scene, args.spp, args.max_depth, args.cpu_budget)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
