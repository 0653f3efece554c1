"""Autotuner for the wavefront's execution knobs.

Searches block shape (threads per traversal block and per forward shade block,
prebuilt kernel variants), wave capacity (path slots per wave) and the grid
sizes of the traversal and shading kernels on a short render of the caller's
workload, and leaves the fastest configuration set on the device. The axes are
tuned one after the other (coordinate descent: block shape, then capacity, then
grids), each with the best values found so far for the others, which costs a
few dozen short renders instead of the full product.

Grid options are multiples of the blocks per CU that are resident at once:
0 is the full grid (one block per queue chunk, balanced by the hardware
dispatcher), k > 0 launches k x resident blocks per CU that stride over the
queue. None of the knobs changes what a path computes: images are the same
for every configuration.
"""
from __future__ import annotations

import time

WAVE_CAPACITIES = (1 << 20, 1 << 21, 1 << 22, 1 << 23, 1 << 24)
TRACE_GRIDS = (0, 1, 2)
SHADE_GRIDS = (0, 1, 2)
TRACE_BLOCKS = (64, 128, 256)
SHADE_BLOCKS = (128, 256, 512)


def tune(device, run, num_cus: int = 256, capacities=WAVE_CAPACITIES, trace_grids=TRACE_GRIDS,
         shade_grids=SHADE_GRIDS, trace_blocks=TRACE_BLOCKS, shade_blocks=SHADE_BLOCKS, min_capacity: int = 0,
         repeats: int = 1, log=None):
    """`run()` performs one short render. Returns (best_config, results)."""
    results = []
    cfg = {"trace_block": 128, "shade_block": 256, "wave_capacity": max(c for c in capacities if c >= min_capacity),
           "trace_grid": 0, "shade_grid": 0}

    def apply(c):
        device.set_block_shape(c["trace_block"], c["shade_block"])
        resident = device.launch_config()
        tb = c["trace_grid"] * max(1, resident["trace_per_cu"]) * num_cus
        sb = c["shade_grid"] * max(1, resident["shade_per_cu"]) * num_cus
        device.set_wave_capacity(c["wave_capacity"])
        device.set_launch_config(tb, sb, True)
        return tb, sb

    def measure(c):
        tb, sb = apply(c)
        run()  # warm-up (graph capture, allocation)
        t0 = time.perf_counter()
        for _ in range(repeats):
            run()
        ms = (time.perf_counter() - t0) * 1000.0 / repeats
        r = dict(c, trace_blocks=tb, shade_blocks=sb, ms=ms)
        results.append(r)
        if log:
            log(f"[autotune] blocks {c['trace_block']}/{c['shade_block']} capacity {c['wave_capacity']} "
                f"grids {tb or 'full'}/{sb or 'full'}: {ms:.2f} ms")
        return ms

    def descend(axes):
        best_ms = None
        best = dict(cfg)
        for values in axes:
            c = dict(cfg, **values)
            ms = measure(c)
            if best_ms is None or ms < best_ms:
                best_ms, best = ms, c
        cfg.update(best)

    descend([{"trace_block": t, "shade_block": s} for t in trace_blocks for s in shade_blocks])
    descend([{"wave_capacity": c} for c in capacities if c >= min_capacity])
    descend([{"trace_grid": t, "shade_grid": s} for t in trace_grids for s in shade_grids])
    apply(cfg)
    best = min((r for r in results if all(r[k] == cfg[k] for k in cfg)), key=lambda r: r["ms"])
    return best, results
