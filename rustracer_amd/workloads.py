"""Seeded synthetic workloads of the benchmark configurations.

multiview_views: 64 cameras at seeded poses around the ~109k-triangle
"multiview" scene, resolutions log-uniform from 128x128 to 1920x1080, spp drawn
uniformly from {1, 2, 4, 8, 16}; all rendered in one persistent launch.
"""
from __future__ import annotations

import numpy as np

from . import Scene


def multiview_views(n: int = 64, seed: int = 1, scale: float = 1.0):
    rng = np.random.default_rng(seed)
    views = []
    for k in range(n):
        u = rng.random()
        w = int(round(128 * (1920 / 128) ** u * scale))
        h = int(round(128 * (1080 / 128) ** u * scale))
        spp = int(rng.choice([1, 2, 4, 8, 16]))
        cam = Scene.builtin_camera("multiview", seed, k)
        views.append((cam, max(w, 1), max(h, 1), spp, k + 1))
    return views
