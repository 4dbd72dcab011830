"""Batched String method for many independent strings on one GPU.

    system = StringSystem(plans, init, directions, precision="mixed")
    system.run(2000, graph=True)
    path, energy, iterations, converged = system.state()

``init`` is (n_strings, n_images, 6): fractional center (a, b, c) and Euler
angles (alpha, beta, gamma) in radians of each image, the reference String
solver's path columns (RC_main.cu:2581-2582 writes a b c alpha beta gamma E).
The iteration follows the reference solver (see csrc/sv_string.hip) with the
analytic gradient in place of forward differences.  ``straight_strings``
builds the straight fractional paths along a, b or c used as starting points.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from .field import POSE_PRECISIONS, Plan, PoseList, _ptr, precision_code

MOVE_FRAC = 1e-4                    # delta_frac of the reference inputs (input.dat)
MOVE_ANGLE = np.radians(1.0)        # delta_angle_degree 1.0


def director(beta, gamma):
    """Body-x guest director after Rz(gamma) Ry(beta) Rx(alpha) (device_fxn.h:34-63)."""
    beta, gamma = np.asarray(beta, float), np.asarray(gamma, float)
    return np.stack([np.cos(gamma) * np.cos(beta), np.sin(gamma) * np.cos(beta), -np.sin(beta)], -1)


class StringSystem:
    def __init__(self, plans, init, directions, precision="fp64", move_frac=MOVE_FRAC,
                 move_angle=MOVE_ANGLE):
        init = np.ascontiguousarray(init, dtype=np.float64)
        if init.ndim != 3 or init.shape[2] != 6:
            raise ValueError("init must be (n_strings, n_images, 6)")
        self.n_strings, self.n_images = init.shape[:2]
        if isinstance(plans, Plan):
            plans = [plans] * self.n_strings
        if len(plans) != self.n_strings:
            raise ValueError("one plan per string required")
        self.plans = list(plans)
        self.lib = self.plans[0].lib
        dirs = np.ascontiguousarray(np.broadcast_to(directions, (self.n_strings,)), dtype=np.int32)
        handles = (ctypes.c_void_p * self.n_strings)(*[p.handle for p in self.plans])
        h = ctypes.c_void_p()
        _lib.check(self.lib.sv_strings_create(self.n_strings, handles, self.n_images, _ptr(init),
                                              _ptr(dirs), float(move_frac), float(move_angle),
                                              precision_code(precision, POSE_PRECISIONS),
                                              ctypes.byref(h)))
        self.handle = h.value

    def run(self, iterations, graph=True):
        _lib.check(self.lib.sv_strings_run(self.handle, int(iterations), int(bool(graph))))

    def state(self):
        S, N = self.n_strings, self.n_images
        path = np.empty((S, N, 6))
        energy = np.empty((S, N))
        its = np.empty(S, dtype=np.int32)
        conv = np.empty(S, dtype=np.int32)
        _lib.check(self.lib.sv_strings_state(self.handle, _ptr(path), _ptr(energy), _ptr(its),
                                             _ptr(conv)))
        return path, energy, its, conv.astype(bool)

    def close(self):
        if getattr(self, "handle", None):
            self.lib.sv_strings_destroy(self.handle)
            self.handle = None

    def __del__(self):
        self.close()


def barrier(energy):
    """Energy span along each path (max - min), as the reference's path audit
    reports it (np.ptp of the path energies)."""
    energy = np.asarray(energy)
    return energy.max(axis=-1) - energy.min(axis=-1)


def straight_strings(plan: Plan, direction: int, n_strings: int, n_images: int, seed: int,
                     scan: int = 24, orientations: int = 8, probes: int = 16,
                     precision: str = "fp64") -> np.ndarray:
    """Straight fractional paths along lattice direction 1/2/3 (a/b/c).

    Each string is the straight line along the hop direction (from 0 to 1) with
    the lowest maximum energy among a seeded scan of scan^2 positions of the
    perpendicular plane x `orientations` fixed orientations, each line probed
    at `probes` points -- the reference solver likewise seeds its string from a
    plane scan (RC_main.cu:622-802) and a straight line (global.h:307-358).
    """
    rng = np.random.default_rng(seed)
    d = direction - 1
    others = [k for k in range(3) if k != d]
    out = np.zeros((n_strings, n_images, 6))
    t = np.linspace(0.0, 1.0, probes)
    for s in range(n_strings):
        grid = rng.random((scan * scan, 2))
        angles = np.c_[np.zeros(orientations), rng.uniform(0, np.pi, orientations),
                       rng.uniform(0, 2 * np.pi, orientations)]
        n_lines = len(grid) * orientations
        frac = np.zeros((n_lines, probes, 3))
        frac[:, :, others[0]] = np.repeat(grid[:, 0], orientations)[:, None]
        frac[:, :, others[1]] = np.repeat(grid[:, 1], orientations)[:, None]
        frac[:, :, d] = t[None, :]
        ang = np.tile(angles, (len(grid), 1))
        dirs = np.repeat(director(ang[:, 1], ang[:, 2]), probes, axis=0)
        e = PoseList(plan, frac.reshape(-1, 3) @ plan.cell, dirs).energy_host(precision)
        best = int(np.argmin(e.reshape(n_lines, probes).max(axis=1)))
        out[s, :, :3] = frac[best, 0]
        out[s, :, d] = np.linspace(0.0, 1.0, n_images)
        out[s, :, 3:] = ang[best]
    return out
