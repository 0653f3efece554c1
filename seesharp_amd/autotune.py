"""Autotuner for the wavefront's execution knobs.

Sweeps the wave capacity (path slots per wave) and the persistent grid sizes of
the traversal and shading kernels on a short render of the caller's workload,
and leaves the fastest configuration set on the device.
"""
from __future__ import annotations

import itertools
import time

WAVE_CAPACITIES = (1 << 18, 1 << 19, 1 << 20, 1 << 21, 1 << 22)
TRACE_BLOCKS_PER_CU = (8, 16, 32)
SHADE_BLOCKS_PER_CU = (4, 8, 16)


def tune(device, run, num_cus: int = 256, capacities=WAVE_CAPACITIES, trace_per_cu=TRACE_BLOCKS_PER_CU,
         shade_per_cu=SHADE_BLOCKS_PER_CU, min_capacity: int = 0, repeats: int = 1, log=None):
    """`run()` performs one short render. Returns (best_config, results)."""
    results = []
    for cap, tb, sb in itertools.product(capacities, trace_per_cu, shade_per_cu):
        if cap < min_capacity:
            continue
        device.set_wave_capacity(cap)
        device.set_launch_config(tb * num_cus, sb * num_cus, True)
        run()  # warm-up (graph capture, allocation)
        t0 = time.perf_counter()
        for _ in range(repeats):
            run()
        ms = (time.perf_counter() - t0) * 1000.0 / repeats
        results.append({"wave_capacity": cap, "trace_blocks": tb * num_cus, "shade_blocks": sb * num_cus, "ms": ms})
        if log:
            log(f"[autotune] capacity {cap} trace {tb}/CU shade {sb}/CU: {ms:.2f} ms")
    best = min(results, key=lambda r: r["ms"])
    device.set_wave_capacity(best["wave_capacity"])
    device.set_launch_config(best["trace_blocks"], best["shade_blocks"], True)
    return best, results
