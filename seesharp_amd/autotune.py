"""Autotuner for the wavefront's execution knobs.

Sweeps the wave capacity (path slots per wave) and the grid sizes of the
traversal and shading kernels on a short render of the caller's workload, and
leaves the fastest configuration set on the device.

Grid options are multiples of the blocks per CU that are resident at once:
0 is the full grid (one block per queue chunk, balanced by the hardware
dispatcher), k > 0 launches k x resident blocks per CU that stride over the
queue.
"""
from __future__ import annotations

import itertools
import time

WAVE_CAPACITIES = (1 << 20, 1 << 21, 1 << 22, 1 << 23, 1 << 24)
TRACE_GRIDS = (0, 1, 2)
SHADE_GRIDS = (0, 1, 2)


def tune(device, run, num_cus: int = 256, capacities=WAVE_CAPACITIES, trace_grids=TRACE_GRIDS,
         shade_grids=SHADE_GRIDS, min_capacity: int = 0, repeats: int = 1, log=None):
    """`run()` performs one short render. Returns (best_config, results)."""
    resident = device.launch_config()
    tpc = max(1, resident["trace_per_cu"])
    spc = max(1, resident["shade_per_cu"])
    results = []
    for cap, tk, sk in itertools.product(capacities, trace_grids, shade_grids):
        if cap < min_capacity:
            continue
        tb, sb = tk * tpc * num_cus, sk * spc * num_cus
        device.set_wave_capacity(cap)
        device.set_launch_config(tb, sb, True)
        run()  # warm-up (graph capture, allocation)
        t0 = time.perf_counter()
        for _ in range(repeats):
            run()
        ms = (time.perf_counter() - t0) * 1000.0 / repeats
        results.append({"wave_capacity": cap, "trace_blocks": tb, "shade_blocks": sb, "ms": ms})
        if log:
            log(f"[autotune] capacity {cap} trace blocks {tb or 'full'} shade blocks {sb or 'full'}: {ms:.2f} ms")
    best = min(results, key=lambda r: r["ms"])
    device.set_wave_capacity(best["wave_capacity"])
    device.set_launch_config(best["trace_blocks"], best["shade_blocks"], True)
    return best, results
