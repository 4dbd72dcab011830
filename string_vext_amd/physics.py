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

import numpy as np

from .orientation import regular_pose_grid


def _finite(a) -> bool:
    return bool(np.isfinite(a).all())


def _positive_vector(values, length) -> bool:
    a = np.asarray(values, dtype=float)
    return a.shape == (length,) and _finite(a) and bool((a > 0).all())


def validate_case(case):
    """Reject a case the exact path cannot evaluate, with the reference's
    messages (physics.py:18-39): the framework (a 3x3 cell of positive
    determinant and a nonempty (n, 3) finite fractional block), its positive LJ
    parameters, a two-site guest with positive parameters and masses laid on the
    body x axis, masses summing to the total, a positive cutoff and temperature."""
    cell = np.asarray(case["cell"], dtype=float)
    frame = np.asarray(case["frame_frac"], dtype=float)
    framework_ok = (cell.shape == (3, 3) and frame.ndim == 2 and frame.shape[1] == 3
                    and len(frame) > 0 and _finite(cell) and _finite(frame)
                    and np.linalg.det(cell) > 0)
    body = np.asarray(case["guest_xyz"], dtype=float)
    checks = (
        (lambda: framework_ok, "invalid triclinic framework"),
        (lambda: all(_positive_vector(case[k], len(frame)) for k in ("frame_sigma", "frame_epsilon")),
         "invalid framework LJ parameters"),
        (lambda: all(_positive_vector(case[k], 2)
                     for k in ("guest_sigma", "guest_epsilon", "guest_mass")),
         "invalid two-site guest"),
        (lambda: body.shape == (2, 3) and _finite(body) and np.allclose(body[:, 1:], 0),
         "guest must have a two-site body x-axis"),
        (lambda: np.isclose(np.sum(case["guest_mass"]), case["total_mass"], rtol=1e-6),
         "guest mass mismatch"),
        (lambda: _finite([case["cutoff"], case["temperature"]])
         and case["cutoff"] > 0 and case["temperature"] > 0,
         "invalid cutoff/temperature"),
    )
    for holds, message in checks:
        if not holds():
            raise ValueError(message)


def unit_directors(centers, axes, message):
    """Validated (n, 3) FP64 centers and row-normalized (k, 3) directors."""
    c = np.asarray(centers, dtype=float)
    u = np.asarray(axes, dtype=float)
    shapes_ok = c.ndim == 2 and c.shape[1:] == (3,) and u.ndim == 2 and u.shape[1:] == (3,)
    if not (shapes_ok and np.isfinite(c).all() and np.isfinite(u).all()):
        raise ValueError(message)
    norm = np.linalg.norm(u, axis=1, keepdims=True)
    if (norm <= 0).any():
        raise ValueError(message)
    return c, np.ascontiguousarray(u / norm)


def body_offsets(case) -> np.ndarray:
    """Signed site offsets along the body axis about the centre of mass (physics.py:88-90)."""
    xyz = np.asarray(case["guest_xyz"], dtype=float)
    mass = np.asarray(case["guest_mass"], dtype=float)
    center = (xyz * mass[:, None]).sum(0) / case["total_mass"]
    return xyz[:, 0] - center[0]


def periodic_images(cell, frac, radius):
    """Framework images reachable within `radius` of any point of the cell
    (physics.py:49-56): with d_k = 1 / |column k of H^-1| the spacing of the
    lattice planes k, whole-cell shifts n_k in [-r_k, r_k], r_k = ceil(radius /
    d_k) + 1 (wrapped query and atom coordinates differ by less than one cell),
    shift-major and atom-minor.  Returns (images [n_shifts * n_atoms][3], n_shifts)."""
    cell = np.asarray(cell, dtype=float)
    plane_spacing = 1.0 / np.linalg.norm(np.linalg.inv(cell), axis=0)
    reach = np.ceil(radius / plane_spacing).astype(int) + 1
    span = 2 * reach + 1
    shifts = (np.indices(span).reshape(3, -1).T - reach).astype(float)
    wrapped = np.mod(np.asarray(frac, dtype=float), 1)
    images = (wrapped[None, :, :] + shifts[:, None, :]).reshape(-1, 3) @ cell
    return images, len(shifts)


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
        self.images, copies = periodic_images(self.cell, case["frame_frac"], self.cutoff)
        self.sigma = np.tile(np.asarray(case["frame_sigma"], dtype=float), copies)
        self.epsilon = np.tile(np.asarray(case["frame_epsilon"], dtype=float), copies)
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
    """Energies [center][axis] of both guest sites at fractional centers and
    (normalized) directors (physics.py:93-106), FP64 on the GPU.  `potential`
    is accepted for signature compatibility; the plan is built from `case`."""
    from .field import Plan
    centers, axes = unit_directors(centers, axes, "invalid center/director queries")
    validate_case(case)
    plan = Plan.from_case(case)
    return plan.field(plan.poses(centers, axes), precision="fp64")["energy"]


def marginal_energy(energies, temperature):
    """Boltzmann marginal over the last (orientation) axis (physics.py:126-130):
    -T ln(mean_k exp(-E_k / T)), evaluated about the lowest energy of each row
    so that no exponential overflows."""
    energy = np.asarray(energies, dtype=float)
    if not (_finite(energy) and np.isfinite(temperature) and temperature > 0):
        raise ValueError("finite angular energies and positive temperature required")
    low = energy.min(axis=-1, keepdims=True)
    mean_weight = np.exp((low - energy) / temperature).mean(axis=-1)
    return low[..., 0] - temperature * np.log(mean_weight)


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
