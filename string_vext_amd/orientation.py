"""Uniform cell-centred pose grid and Fibonacci directions.

Mirrors ``string_vext.orientation.regular_pose_grid`` (reference
src/string_vext/orientation.py:7-16): centers (i + 0.5)/n on an ij-indexed
meshgrid, flat index (ia n + ib) n + ic; direction k has z = 1 - 2 (k + 0.5)/K
and azimuth k pi (3 - sqrt 5).
"""
from __future__ import annotations

import numpy as np


def regular_pose_grid(size: int, orientations: int):
    if size < 2 or orientations < 4:
        raise ValueError("at least two grid cells and four orientations required")
    ticks = (np.arange(size, dtype=float) + 0.5) / size
    mesh = np.meshgrid(ticks, ticks, ticks, indexing="ij")
    centers = np.stack(mesh, axis=-1).reshape(-1, 3)
    k = np.arange(orientations)
    z = 1 - 2 * (k + 0.5) / orientations
    phi = k * np.pi * (3 - np.sqrt(5))
    rho = np.sqrt(1 - z * z)
    return centers, np.stack([rho * np.cos(phi), rho * np.sin(phi), z], axis=1)
