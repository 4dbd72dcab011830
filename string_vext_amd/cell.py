"""Triclinic cell matrix from lattice lengths and angles.

Mirrors ``string_vext.cell.cell_matrix`` (reference src/string_vext/cell.py:7-29):
rows are the lattice vectors a = (a, 0, 0), b = (b cos g, b sin g, 0),
c = (c cos b, c (cos a - cos b cos g) / sin g, sqrt(c^2 - cx^2 - cy^2)).
Host-side FP64 setup; the GPU path receives the matrix.
"""
from __future__ import annotations

import numpy as np


def cell_matrix(lengths, angles_deg) -> np.ndarray:
    lengths = np.asarray(lengths, dtype=float)
    angles = np.asarray(angles_deg, dtype=float)
    if (lengths.shape != (3,) or angles.shape != (3,) or not np.isfinite(lengths).all()
            or not np.isfinite(angles).all() or (lengths <= 0).any() or (angles <= 0).any()
            or (angles >= 180).any()):
        raise ValueError("three positive lengths and three angles in (0, 180) required")
    alpha, beta, gamma = np.deg2rad(angles)
    a, b, c = lengths
    sin_g = np.sin(gamma)
    if abs(sin_g) < 1e-10:
        raise ValueError("singular triclinic cell")
    cx = c * np.cos(beta)
    cy = c * (np.cos(alpha) - np.cos(beta) * np.cos(gamma)) / sin_g
    cz2 = c * c - cx * cx - cy * cy
    if cz2 <= 0:
        raise ValueError("nonphysical triclinic cell")
    m = np.array([[a, 0.0, 0.0], [b * np.cos(gamma), b * sin_g, 0.0], [cx, cy, np.sqrt(cz2)]])
    if not np.isfinite(m).all() or np.linalg.det(m) <= 0:
        raise ValueError("invalid cell determinant")
    return m
