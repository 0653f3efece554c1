"""Writes tests/golden/large_scene_ref_4096spp.npz: the oracle's converged render of a
seeded 8192-pixel subset of the large scene (1 036 296 triangles, 1024x1024, MaxDepth 8)
at 4096 spp with BaseSeed 571298512 (the reference's own reference-render seed,
SeeSharp/Experiments/SceneFromFile.cs:191), independent of the measured runs' seed.

    python tests/golden/make_converged_reference.py
"""
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import oracle_lib as ol  # noqa: E402
from seesharp_amd import scenes  # noqa: E402

SPP = 4096
NUM_PIXELS = 8192
SEED = 571298512


def main():
    scene = scenes.large_scene(1024, 1024)
    o = ol.OracleScene(scene)
    pix = np.sort(np.random.default_rng(20261203).choice(1024 * 1024, NUM_PIXELS, replace=False)).astype(np.int32)
    cfg = ol.settings(base_seed=SEED, total_spp=SPP, max_depth=8)
    img = np.zeros((1024, 1024, 3), np.float32)
    sq = np.zeros((NUM_PIXELS, 3), np.float64)
    t0 = time.time()
    # 64 chunks of 64 iterations; per-chunk means give a standard error per pixel
    chunk_means = []
    for c in range(SPP // 64):
        part = np.zeros((1024, 1024, 3), np.float32)
        o.render(cfg, c * 64, 64, image=part, pixels=pix)
        # part holds the running mean of iterations c*64..c*64+63 started at iteration c*64:
        # its scale factors assume earlier iterations, so undo them to get the chunk mean
        n0, n1 = c * 64, c * 64 + 64
        chunk = part.reshape(-1, 3)[pix] * (n1 / (n1 - n0))
        chunk_means.append(chunk.astype(np.float64))
        print(f"chunk {c + 1}/{SPP // 64} {time.time() - t0:.0f}s", flush=True)
    cm = np.stack(chunk_means)
    mean = cm.mean(0)
    se = cm.std(0, ddof=1) / np.sqrt(cm.shape[0])
    np.savez_compressed(os.path.join(HERE, "large_scene_ref_4096spp.npz"), pixels=pix,
                        mean=mean.astype(np.float32), stderr=se.astype(np.float32), spp=SPP, base_seed=SEED)
    print("done", time.time() - t0)


if __name__ == "__main__":
    main()
