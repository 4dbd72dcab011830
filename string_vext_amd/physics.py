"""Full triclinic image sums for String LJ parameters: validation, images, site
field, DIRECT pose energies and the orientational marginal.

Mirrors ``string_vext.physics`` (reference src/string_vext/physics.py).  The
energies (``TriclinicSitePotential.evaluate``, ``direct_pose_energy``,
``uniform_fields``) are computed on the GPU by libsv_vext.so; this module only
prepares FP64 geometry on the host.  The interpolation approximation of the
reference (``periodic_interpolate``, ``energy_mode='interpolate'``) is not part
of the exact path and is not provided.
"""
from __future__ import annotations

import itertools

import numpy as np

from .orientation import regular_pose_grid


def validate_case(case):
    """Reference physics.py:18-39."""
    cell = np.asarray(case["cell"], dtype=float)
    frame = np.asarray(case["frame_frac"], dtype=float)
    if (cell.shape != (3, 3) or not np.isfinite(cell).all() or np.linalg.det(cell) <= 0
            or frame.ndim != 2 or frame.shape[1] != 3 or not len(frame)
            or not np.isfinite(frame).all()):
        raise ValueError("invalid triclinic framework")
    for key in ("frame_sigma", "frame_epsilon"):
        values = np.asarray(case[key], dtype=float)
        if values.shape != (len(frame),) or not np.isfinite(values).all() or (values <= 0).any():
            raise ValueError("invalid framework LJ parameters")
    for key in ("guest_sigma", "guest_epsilon", "guest_mass"):
        values = np.asarray(case[key], dtype=float)
        if values.shape != (2,) or not np.isfinite(values).all() or (values <= 0).any():
            raise ValueError("invalid two-site guest")
    body = np.asarray(case["guest_xyz"], dtype=float)
    if body.shape != (2, 3) or not np.isfinite(body).all() or not np.allclose(body[:, 1:], 0):
        raise ValueError("guest must have a two-site body x-axis")
    if not np.isclose(sum(case["guest_mass"]), case["total_mass"], rtol=1e-6):
        raise ValueError("guest mass mismatch")
    if (not np.isfinite([case["cutoff"], case["temperature"]]).all()
            or min(case["cutoff"], case["temperature"]) <= 0):
        raise ValueError("invalid cutoff/temperature")


def body_offsets(case) -> np.ndarray:
    """Signed site offsets along the body axis about the centre of mass (physics.py:88-90)."""
    xyz = np.asarray(case["guest_xyz"], dtype=float)
    mass = np.asarray(case["guest_mass"], dtype=float)
    center = (xyz * mass[:, None]).sum(0) / case["total_mass"]
    return xyz[:, 0] - center[0]


class TriclinicSitePotential:
    """Framework images within the cutoff of any point of the cell (physics.py:42-57).

    ``images``/``sigma``/``epsilon`` are the FP64 image table the drop-in C ABI
    consumes; ``tree`` (a scipy cKDTree over the images, built on first use) is
    the candidate filter of that ABI's CSR.  ``evaluate`` runs on the GPU.
    """

    def __init__(self, case):
        validate_case(case)
        self.case = case
        self.cell = np.asarray(case["cell"], dtype=np.float64)
        self.inverse = np.linalg.inv(self.cell)
        self.cutoff = float(case["cutoff"])
        spacing = 1 / np.linalg.norm(self.inverse, axis=0)
        # wrapped fractional query/atom components differ by less than one cell
        reach = np.ceil(self.cutoff / spacing).astype(int) + 1
        shifts = np.array(list(itertools.product(*[range(-r, r + 1) for r in reach])), dtype=float)
        atoms = np.mod(case["frame_frac"], 1)
        self.images = (atoms[None] + shifts[:, None]).reshape(-1, 3) @ self.cell
        self.sigma = np.tile(case["frame_sigma"], len(shifts))
        self.epsilon = np.tile(case["frame_epsilon"], len(shifts))
        self._tree = None

    @property
    def tree(self):
        if self._tree is None:
            from scipy.spatial import cKDTree
            self._tree = cKDTree(self.images)
        return self._tree

    def evaluate(self, coordinates, guest_sigma, guest_epsilon, chunk=None):
        """Single-site energy at Cartesian points, FP64 on the GPU.

        Evaluated as a two-site pose whose sites coincide (offsets 0, 0), halved:
        the same shifted, floored LJ sum over every image within the cutoff.
        """
        from .field import Plan
        query = np.asarray(coordinates, dtype=float)
        if query.ndim != 2 or query.shape[1] != 3 or not np.isfinite(query).all():
            raise ValueError("invalid Cartesian site coordinates")
        case = self.case
        plan = Plan(self.cell, case["frame_frac"], case["frame_sigma"], case["frame_epsilon"],
                    guest_sigma, guest_epsilon, (0.0, 0.0), self.cutoff)
        poses = plan.poses(query @ self.inverse, np.array([[1.0, 0.0, 0.0]]))
        energy = plan.field(poses, precision="fp64")["energy"][:, 0] / 2.0
        if not np.isfinite(energy).all():
            raise ValueError("nonfinite String-FF site energy")
        return energy


def direct_pose_energy(case, centers, axes, potential=None):
    """Energies [center][axis] of both guest sites (physics.py:93-106), GPU FP64."""
    from .field import Plan
    centers, axes = np.asarray(centers, dtype=float), np.asarray(axes, dtype=float)
    if (centers.ndim != 2 or centers.shape[1] != 3 or axes.ndim != 2 or axes.shape[1] != 3
            or not np.isfinite(centers).all() or not np.isfinite(axes).all()
            or (np.linalg.norm(axes, axis=1) <= 0).any()):
        raise ValueError("invalid center/director queries")
    validate_case(case)
    axes = axes / np.linalg.norm(axes, axis=1, keepdims=True)
    plan = Plan.from_case(case)
    return plan.field(plan.poses(centers, axes), precision="fp64")["energy"]


def marginal_energy(energies, temperature):
    """Boltzmann marginal over the last (orientation) axis (physics.py:126-130)."""
    from scipy.special import logsumexp
    energy = np.asarray(energies, dtype=float)
    if not np.isfinite(energy).all() or not np.isfinite(temperature) or temperature <= 0:
        raise ValueError("finite angular energies and positive temperature required")
    return -temperature * (logsumexp(-energy / temperature, axis=-1) - np.log(energy.shape[-1]))


def uniform_fields(case, size=60, orientations=64, energy_mode="direct", precision="fp64"):
    """Orientation channels, marginal, site fields and axes (physics.py:133-164), GPU."""
    if energy_mode != "direct":
        raise NotImplementedError("only the exact 'direct' mode is provided")
    from .field import Plan
    validate_case(case)
    centers, axes = regular_pose_grid(size, orientations)
    plan = Plan.from_case(case)
    out = plan.field(plan.grid(size, axes), precision=precision, energy=False,
                     orientation=True, marginal=True, site=True,
                     temperature=case["temperature"])
    site = out["site"].reshape(size, size, size)
    return dict(orientation_K=out["orientation"].reshape(orientations, size, size, size),
                marginal_K=out["marginal"].reshape(size, size, size),
                site_K=np.stack([site, site]).astype(np.float32), axes=axes.astype(np.float32))
