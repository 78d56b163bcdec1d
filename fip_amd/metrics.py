"""Error measures of a coarse run against the restricted fine run of the same scenario
(the measures the learned closure is scored by): depth RMSE, max-depth error, wet-extent
IoU."""
from __future__ import annotations

import numpy as np


def depth_rmse(h, h_ref) -> float:
    d = np.asarray(h, np.float64) - np.asarray(h_ref, np.float64)
    return float(np.sqrt(np.mean(d * d)))


def max_depth_error(hmax, hmax_ref) -> float:
    return float(np.max(np.abs(np.asarray(hmax, np.float64) - np.asarray(hmax_ref, np.float64))))


def wet_iou(h, h_ref, thresh: float = 0.01) -> float:
    """Intersection over union of the cells deeper than `thresh` (m)."""
    a = np.asarray(h) > thresh
    b = np.asarray(h_ref) > thresh
    union = np.logical_or(a, b).sum()
    return float(np.logical_and(a, b).sum() / union) if union else 1.0
