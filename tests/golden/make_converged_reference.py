"""Writes tests/golden/large_scene_ref_4096spp.npz: the oracle's converged render of a
seeded 8192-pixel subset of the large scene (1 036 296 triangles, 1024x1024, MaxDepth 8)
at 4096 spp with BaseSeed 571298512 (the reference's own reference-render seed,
SeeSharp/Experiments/SceneFromFile.cs:191), independent of the measured runs' seed,
and the same pixels rendered with the default BaseSeed 0xC030114 (the GPU runs' seed):
the GPU must match the latter path for path, and both CPU renders together give the
Monte Carlo noise floor of the relMSE between two independent 4096-spp images.

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


def render(o, pix, seed):
    cfg = ol.settings(base_seed=seed, total_spp=SPP, max_depth=8)
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
    return cm.mean(0), cm.std(0, ddof=1) / np.sqrt(cm.shape[0])


def main():
    scene = scenes.large_scene(1024, 1024)
    o = ol.OracleScene(scene)
    pix = np.sort(np.random.default_rng(20261203).choice(1024 * 1024, NUM_PIXELS, replace=False)).astype(np.int32)
    mean, se = render(o, pix, SEED)
    same, _ = render(o, pix, 0xC030114)
    np.savez_compressed(os.path.join(HERE, "large_scene_ref_4096spp.npz"), pixels=pix,
                        mean=mean.astype(np.float32), stderr=se.astype(np.float32), spp=SPP, base_seed=SEED,
                        mean_same_seed=same.astype(np.float32), same_seed=0xC030114)


if __name__ == "__main__":
    main()
