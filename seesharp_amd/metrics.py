"""Image error metrics of the FrameBuffer's per-iteration error log
(FrameBuffer.ComputeErrorMetric, SeeSharp/Images/FrameBuffer.cs:377-383).

The reference takes them from SimpleImageIO's `Metrics`, whose source is not
part of the reference; the definitions here are this package's own and are
stated exactly: over all pixels and channels, in float64,

* MSE = mean((x - r)^2)
* RelMSE = mean((x - r)^2 / (r^2 + epsilon)), epsilon = 0.01
* RelMSE_OutlierRejection: the per-pixel relative squared error (averaged over
  the channels), the largest `percentage` percent of pixels dropped, the mean
  of the rest (percentage = 0.1, i.e. 0.1 % of the pixels).
"""
from __future__ import annotations

import numpy as np


def _pair(image, reference):
    x = np.asarray(image, np.float64)
    r = np.asarray(reference, np.float64)
    if x.shape != r.shape:
        raise ValueError(f"image {x.shape} and reference {r.shape} differ in size")
    return x, r


def MSE(image, reference) -> float:
    x, r = _pair(image, reference)
    return float(np.mean((x - r) ** 2))


def RelMSE(image, reference, epsilon: float = 0.01) -> float:
    x, r = _pair(image, reference)
    return float(np.mean((x - r) ** 2 / (r * r + epsilon)))


def RelMSE_OutlierRejection(image, reference, epsilon: float = 0.01, percentage: float = 0.1) -> float:
    x, r = _pair(image, reference)
    per_pixel = ((x - r) ** 2 / (r * r + epsilon)).reshape(x.shape[0] * x.shape[1], -1).mean(axis=1)
    keep = per_pixel.size - int(per_pixel.size * percentage / 100.0)
    return float(np.sort(per_pixel)[:max(1, keep)].mean())


def ErrorMetric(render_time_ms: float, image, reference) -> dict:
    """One entry of FrameBuffer.Errors (FrameBuffer.cs:377-383)."""
    return {"RenderTime": render_time_ms, "MSE": MSE(image, reference), "RelMSE": RelMSE(image, reference),
            "RelMSE_OutlierRejection": RelMSE_OutlierRejection(image, reference)}
