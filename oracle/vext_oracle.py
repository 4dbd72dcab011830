"""FP64 CPU oracle of the guest-framework LJ Vext path.

TEST INFRASTRUCTURE ONLY: imported by ``tests/``, ``__graft_entry__.smoke()``
and the ``cpu_baseline`` leg of ``bench.py`` -- as the checker, or as the
timed CPU baseline -- and never by the ``string_vext_amd`` package.

Every function restates the reference at the cited file:line (paths are
relative to the reference repository).  The arithmetic of the pose sum lives
in ``vext_oracle.c`` (``oracle_reuse``); this module restates the host-side
preparation the reference does in numpy/scipy.

How each part is pinned (tests/test_oracle.py):
  * ``oracle_reuse`` equals ``oracle/_ref/kernel.so`` -- the reference's own
    ``native/vext/kernel.cpp`` compiled from source -- bit for bit on the same
    CSR, and a brute-force sum over an oversized block of lattice images;
  * images / coefficients / candidate sets: completeness against the
    brute-force enumeration, closed forms of the mixing rules;
  * ``cell_matrix``: the KAT of tests/unit/test_public_api.py:43-50 and
    lattice-vector lengths/angles recovered from the matrix;
  * ``regular_pose_grid``: closed-form Fibonacci directions (unit norm,
    z-levels, golden-angle increments);
  * ``direct_pose_energy`` vs ``reuse_energy`` at the tolerances of
    tests/unit/test_public_api.py:60-61;
  * ``string_pair`` equals ``oracle/_ref/string_pair.so`` (the reference's
    triclinic_pair.h through its own host shim) and the brute force of
    test_native_string_triclinic.py:41-49;
  * ``pose_grad`` against central finite differences of ``pose_grad``'s own
    energy (gradient parity), the energy itself against ``oracle_reuse``;
  * ``parse_string_input`` (tests/test_oracle_pins.py): the literal values of
    the reference's toy ``input.dat`` and of tests/cof/examples/c2h4/input.dat
    (kept under tests/golden/), the cell KAT through the parser, a planted
    column swap must fail;
  * ``body_offsets``: closed form for site masses 10 and 20, and
    sum(m_s o_s) = 0, o_1 - o_0 = x_1 - x_0 for random unequal masses;
  * ``count_pairs`` (the metric's numerator): a brute-force count over an
    oversized image block, and a pair exactly at the cutoff not counted;
  * ``marginal_energy``: closed forms (degenerate and two-level);
  * ``error_metrics``: the accessible mask and the 1 K floor on constructed
    inputs.
No function here is left unpinned.
"""
from __future__ import annotations

import ctypes
import itertools
import subprocess
from pathlib import Path

import numpy as np
from scipy.spatial import cKDTree
from scipy.special import logsumexp

HERE = Path(__file__).resolve().parent
ORACLE_LIB = HERE / "liboracle_vext.so"
REF_KERNEL = HERE / "_ref" / "kernel.so"
REF_STRING_PAIR = HERE / "_ref" / "string_pair.so"

_P = ctypes.c_void_p
_lib = None


def build() -> None:
    """Compile liboracle_vext.so (and oracle/_ref when the reference is present)."""
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(_P)


def lib():
    global _lib
    if _lib is None:
        if not ORACLE_LIB.exists():
            build()
        handle = ctypes.CDLL(str(ORACLE_LIB))
        handle.oracle_reuse.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                        ctypes.c_double] + [_P] * 11
        handle.oracle_reuse.restype = ctypes.c_int
        handle.oracle_site.argtypes = [ctypes.c_int, _P, _P, _P, _P, _P, ctypes.c_double,
                                       _P, _P, _P]
        handle.oracle_site.restype = ctypes.c_int
        handle.oracle_pose_grad.argtypes = [ctypes.c_int] + [_P] * 8 + [ctypes.c_double] \
            + [_P] * 5
        handle.oracle_pose_grad.restype = ctypes.c_int
        handle.oracle_string_pair.argtypes = [ctypes.c_double] * 6 + [_P] * 3 \
            + [ctypes.c_double] * 5
        handle.oracle_string_pair.restype = ctypes.c_double
        handle.oracle_count_pairs.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_double] \
            + [_P] * 6
        handle.oracle_count_pairs.restype = ctypes.c_int64
        _lib = handle
    return _lib


def reference_kernel():
    """The reference's evaluate_reused_neighbors built from its own source, or None."""
    if not REF_KERNEL.exists():
        return None
    handle = ctypes.CDLL(str(REF_KERNEL))
    fn = handle.evaluate_reused_neighbors
    fn.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double] + [_P] * 11
    fn.restype = ctypes.c_int
    fn._keep = handle
    return fn


def reference_string_pair():
    """run_triclinic_pair of the reference's host shim, or None."""
    if not REF_STRING_PAIR.exists():
        return None
    handle = ctypes.CDLL(str(REF_STRING_PAIR))
    fn = handle.run_triclinic_pair
    fn.argtypes = [ctypes.POINTER(ctypes.c_double)]
    fn.restype = ctypes.c_double
    fn._keep = handle
    return fn


# ---------------------------------------------------------------------------
# Host-side preparation (restated from the reference's numpy code)
# ---------------------------------------------------------------------------
def cell_matrix(lengths, angles_deg) -> np.ndarray:
    """src/string_vext/cell.py:7-29 -- rows a, b, c of the triclinic cell."""
    a, b, c = (float(v) for v in lengths)
    al, be, ga = np.deg2rad(np.asarray(angles_deg, dtype=float))
    cx = c * np.cos(be)
    cy = c * (np.cos(al) - np.cos(be) * np.cos(ga)) / np.sin(ga)
    cz2 = c * c - cx * cx - cy * cy
    if cz2 <= 0:
        raise ValueError("nonphysical triclinic cell")
    return np.array([[a, 0.0, 0.0],
                     [b * np.cos(ga), b * np.sin(ga), 0.0],
                     [cx, cy, np.sqrt(cz2)]])


def regular_pose_grid(size: int, orientations: int):
    """src/string_vext/orientation.py:7-16 -- cell-centred grid + Fibonacci directions."""
    ticks = (np.arange(size, dtype=float) + 0.5) / size
    ia, ib, ic = np.meshgrid(ticks, ticks, ticks, indexing="ij")
    centers = np.stack([ia.ravel(), ib.ravel(), ic.ravel()], axis=1)
    k = np.arange(orientations)
    z = 1.0 - 2.0 * (k + 0.5) / orientations
    phi = k * np.pi * (3.0 - np.sqrt(5.0))
    rho = np.sqrt(1.0 - z * z)
    return centers, np.stack([rho * np.cos(phi), rho * np.sin(phi), z], axis=1)


def parse_string_input(path) -> dict:
    """src/string_vext/input.py:11-57 -- the native String input.dat format."""
    lines = [line.strip() for line in Path(path).read_text().splitlines()]

    def row_after(marker):
        for i, line in enumerate(lines):
            if line.startswith(marker):
                return i + 1
        raise ValueError(f"missing {marker!r}")

    lengths = [float(v) for v in lines[row_after("La Lb Lc")].split()[:3]]
    angles = [float(v) for v in lines[row_after("Alpha Beta Gamma")].split()[:3]]
    cond = [float(v) for v in lines[row_after("cutoff(A)")].split()]
    g0 = row_after("Number of sites")
    ng = int(lines[g0])
    guest = np.array([[float(v) for v in lines[g0 + 2 + i].split()[:7]] for i in range(ng)])
    f0 = row_after("Number of atoms")
    nf = int(lines[f0])
    frame = np.array([[float(v) for v in lines[f0 + 2 + i].split()[:8]] for i in range(nf)])
    mass = guest[:, 6]
    body = guest[:, :3]
    sep = body[1] - body[0]
    bond = float(np.linalg.norm(sep))
    director = sep / bond
    com = (body * mass[:, None]).sum(0) / cond[2]
    xyz = np.zeros((2, 3))
    xyz[:, 0] = (body - com) @ director
    return {"cell": cell_matrix(lengths, angles), "frame_frac": np.mod(frame[:, 5:8], 1),
            "frame_sigma": frame[:, 1], "frame_epsilon": frame[:, 2], "guest_xyz": xyz,
            "guest_sigma": guest[:, 4], "guest_epsilon": guest[:, 3], "guest_mass": mass,
            "total_mass": float(cond[2]), "temperature": float(cond[3]),
            "cutoff": float(cond[0]), "cell_lengths": lengths, "cell_angles_deg": angles,
            "bond_length": bond}


def body_offsets(case) -> np.ndarray:
    """src/string_vext/physics.py:88-90 -- site offsets along the body x-axis."""
    xyz = np.asarray(case["guest_xyz"], dtype=float)
    mass = np.asarray(case["guest_mass"], dtype=float)
    com_x = (xyz[:, 0] * mass).sum() / case["total_mass"]
    return xyz[:, 0] - com_x


def lattice_images(cell, frac, radius):
    """src/string_vext/physics.py:49-56 -- every image a query inside the cell can
    reach within `radius`: reach_k = ceil(radius / d_k) + 1 with d_k the
    spacing between lattice planes k, atoms wrapped to [0,1)."""
    cell = np.asarray(cell, dtype=float)
    inv = np.linalg.inv(cell)
    spacing = 1.0 / np.linalg.norm(inv, axis=0)
    reach = np.ceil(radius / spacing).astype(int) + 1
    shifts = np.array(list(itertools.product(*[range(-r, r + 1) for r in reach])), dtype=float)
    atoms = np.mod(np.asarray(frac, dtype=float), 1)
    return (atoms[None] + shifts[:, None]).reshape(-1, 3) @ cell, len(shifts)


class ReuseSetup:
    """src/string_vext/neighbor.py:9-30 -- images within reach of R = rc + max|o|
    and their mixed LJ coefficients, plus the cKDTree used for candidates."""

    def __init__(self, case):
        sig = np.asarray(case["guest_sigma"], dtype=float)
        eps = np.asarray(case["guest_epsilon"], dtype=float)
        if sig.shape != (2,) or eps.shape != (2,) or sig[0] != sig[1] or eps[0] != eps[1]:
            raise ValueError("identical two-site guest required")
        self.offsets = body_offsets(case)
        self.cutoff = float(case["cutoff"])
        self.radius = self.cutoff + float(np.abs(self.offsets).max())
        self.cell = np.asarray(case["cell"], dtype=float)
        self.images, nshift = lattice_images(self.cell, case["frame_frac"], self.radius)
        fs = np.tile(np.asarray(case["frame_sigma"], dtype=float), nshift)
        fe = np.tile(np.asarray(case["frame_epsilon"], dtype=float), nshift)
        ms = (fs + sig[0]) / 2
        me = np.sqrt(fe * eps[0])
        s6 = ms ** 6
        self.coeff12 = 4 * me * s6 ** 2
        self.coeff6 = 4 * me * s6
        c6 = (ms / self.cutoff) ** 6
        self.shift = -4 * me * (c6 ** 2 - c6)
        self.floor2 = (0.1 * ms) ** 2
        self.tree = cKDTree(self.images)


def csr(tree, points, radius):
    """src/string_vext/native.py:47-55 -- candidate CSR from cKDTree ball queries."""
    groups = tree.query_ball_point(points, radius, workers=1)
    lengths = np.fromiter((len(g) for g in groups), dtype=np.int64, count=len(groups))
    starts = np.zeros(len(groups) + 1, dtype=np.int64)
    starts[1:] = np.cumsum(lengths)
    if starts[-1]:
        indices = np.concatenate([np.asarray(g, dtype=np.int32) for g in groups])
    else:
        indices = np.empty(0, np.int32)
    return starts, indices


def run_reuse_kernel(fn, centers_cart, axes, setup: ReuseSetup, starts, indices) -> np.ndarray:
    """Call an evaluate_reused_neighbors-shaped function (oracle or reference)."""
    centers_cart = np.ascontiguousarray(centers_cart, dtype=np.float64)
    axes = np.ascontiguousarray(axes, dtype=np.float64)
    out = np.empty((len(centers_cart), len(axes)), dtype=np.float64)
    arrays = [centers_cart, axes, np.ascontiguousarray(setup.images), np.ascontiguousarray(setup.offsets),
              np.ascontiguousarray(setup.coeff12), np.ascontiguousarray(setup.coeff6),
              np.ascontiguousarray(setup.shift), np.ascontiguousarray(setup.floor2),
              np.ascontiguousarray(starts, dtype=np.int64),
              np.ascontiguousarray(indices, dtype=np.int32), out]
    rc = fn(len(centers_cart), len(axes), len(setup.images), setup.cutoff ** 2,
            *[_ptr(a) for a in arrays])
    if rc:
        raise RuntimeError(f"kernel failed with code {rc}")
    return out


def reuse_energy(case, centers_frac, axes, kernel="oracle", setup=None, batch=512):
    """src/string_vext/native.py:26-68 -- pose energies [center][axis] in FP64."""
    setup = ReuseSetup(case) if setup is None else setup
    centers_frac = np.asarray(centers_frac, dtype=float)
    axes = np.asarray(axes, dtype=float)
    axes = axes / np.linalg.norm(axes, axis=1, keepdims=True)
    cart = np.mod(centers_frac, 1) @ setup.cell
    fn = lib().oracle_reuse if kernel == "oracle" else reference_kernel()
    if fn is None:
        raise RuntimeError("oracle/_ref/kernel.so is not built")
    out = np.empty((len(cart), len(axes)))
    for start in range(0, len(cart), batch):
        chunk = cart[start:start + batch]
        starts, indices = csr(setup.tree, chunk, setup.radius)
        out[start:start + batch] = run_reuse_kernel(fn, chunk, axes, setup, starts, indices)
    return out


def count_pairs(case, centers_frac, axes, setup=None, batch=2048) -> int:
    """Exact number of in-cutoff (site, image) pairs over all (center, axis) poses."""
    setup = ReuseSetup(case) if setup is None else setup
    axes = np.asarray(axes, dtype=float)
    axes = np.ascontiguousarray(axes / np.linalg.norm(axes, axis=1, keepdims=True))
    cart = np.mod(np.asarray(centers_frac, dtype=float), 1) @ setup.cell
    total = 0
    images = np.ascontiguousarray(setup.images)
    offsets = np.ascontiguousarray(setup.offsets)
    for start in range(0, len(cart), batch):
        chunk = np.ascontiguousarray(cart[start:start + batch])
        starts, indices = csr(setup.tree, chunk, setup.radius)
        total += lib().oracle_count_pairs(len(chunk), len(axes), setup.cutoff ** 2, _ptr(chunk),
                                          _ptr(axes), _ptr(images), _ptr(offsets), _ptr(starts),
                                          _ptr(indices))
    return int(total)


def site_energy(case, points_cart, guest_sigma, guest_epsilon):
    """src/string_vext/physics.py:59-85 -- single-site field with radius rc."""
    cutoff = float(case["cutoff"])
    cell = np.asarray(case["cell"], dtype=float)
    images, nshift = lattice_images(cell, case["frame_frac"], cutoff)
    fs = np.tile(np.asarray(case["frame_sigma"], dtype=float), nshift)
    fe = np.tile(np.asarray(case["frame_epsilon"], dtype=float), nshift)
    ms = (fs + guest_sigma) / 2
    me = np.sqrt(fe * guest_epsilon)
    rcut = (ms / cutoff) ** 6
    shifted = -4 * me * (rcut ** 2 - rcut)
    query = np.mod(np.asarray(points_cart, dtype=float) @ np.linalg.inv(cell), 1) @ cell
    starts, indices = csr(cKDTree(images), query, cutoff)
    out = np.empty(len(query))
    arrays = [np.ascontiguousarray(query), np.ascontiguousarray(images), np.ascontiguousarray(ms),
              np.ascontiguousarray(me), np.ascontiguousarray(shifted)]
    lib().oracle_site(len(query), *[_ptr(a) for a in arrays], cutoff, _ptr(starts),
                      _ptr(indices), _ptr(out))
    return out


def direct_pose_energy(case, centers_frac, axes):
    """src/string_vext/physics.py:93-106 -- DIRECT: site fields at both sites."""
    axes = np.asarray(axes, dtype=float)
    axes = axes / np.linalg.norm(axes, axis=1, keepdims=True)
    cell = np.asarray(case["cell"], dtype=float)
    centers = np.asarray(centers_frac, dtype=float)
    energy = np.zeros((len(centers), len(axes)))
    for site, offset in enumerate(body_offsets(case)):
        pos = centers[:, None] @ cell + offset * axes[None]
        energy += site_energy(case, pos.reshape(-1, 3), case["guest_sigma"][site],
                              case["guest_epsilon"][site]).reshape(energy.shape)
    return energy


def marginal_energy(energy, temperature):
    """src/string_vext/physics.py:126-130 -- Boltzmann marginal over orientations."""
    energy = np.asarray(energy, dtype=float)
    return -temperature * (logsumexp(-energy / temperature, axis=-1) - np.log(energy.shape[-1]))


def vext_fields(case, size, orientations, setup=None):
    """src/string_vext/vext.py:21-43 over the oracle: every NPZ channel."""
    centers, axes = regular_pose_grid(size, orientations)
    energy = reuse_energy(case, centers, axes, setup=setup)
    marginal = marginal_energy(energy, case["temperature"])
    site = site_energy(case, centers @ np.asarray(case["cell"]), case["guest_sigma"][0],
                       case["guest_epsilon"][0])
    out = {"energy": energy,
           "orientation_K": energy.T.reshape(orientations, size, size, size).astype(np.float32),
           "marginal_K": marginal.reshape(size, size, size).astype(np.float32),
           "site_K": np.stack((site, site)).reshape(2, size, size, size).astype(np.float32),
           "axes": axes.astype(np.float32)}
    out["grid"] = np.arcsinh(out["marginal_K"] / 1000).astype(np.float32)
    return out


def brute_pose_energy(case, centers_frac, axes, reach=3):
    """Pose energies from an oversized block of lattice shifts (-reach..reach)^3,
    no neighbor list: the independent check of image completeness (the same
    construction as test_native_string_triclinic.py:41-49).  Tiny inputs only."""
    cell = np.asarray(case["cell"], dtype=float)
    shifts = np.array(list(itertools.product(range(-reach, reach + 1), repeat=3)), dtype=float)
    atoms = np.mod(np.asarray(case["frame_frac"], dtype=float), 1)
    images = ((atoms[None] + shifts[:, None]).reshape(-1, 3)) @ cell
    n = len(shifts)
    fs = np.tile(np.asarray(case["frame_sigma"], dtype=float), n)
    fe = np.tile(np.asarray(case["frame_epsilon"], dtype=float), n)
    ms = (fs + case["guest_sigma"][0]) / 2
    me = np.sqrt(fe * case["guest_epsilon"][0])
    rc = float(case["cutoff"])
    c6 = (ms / rc) ** 6
    axes = np.asarray(axes, dtype=float)
    axes = axes / np.linalg.norm(axes, axis=1, keepdims=True)
    cart = np.mod(np.asarray(centers_frac, dtype=float), 1) @ cell
    out = np.zeros((len(cart), len(axes)))
    for s in body_offsets(case):
        sites = cart[:, None, :] + s * axes[None, :, :]
        d2 = ((sites[:, :, None, :] - images[None, None]) ** 2).sum(-1)
        q = np.maximum(d2, (0.1 * ms) ** 2)
        r6 = (ms ** 2 / q) ** 3
        term = 4 * me * (r6 * r6 - r6) - 4 * me * (c6 * c6 - c6)
        out += np.where(d2 < rc * rc, term, 0.0).sum(-1)
    return out


def pose_grad(case, centers_cart, axes, setup=None):
    """FP64 energy and analytic gradient of independent poses (center_i, axis_i),
    candidates within R of each center; axes used as given (unit)."""
    setup = ReuseSetup(case) if setup is None else setup
    cell = setup.cell
    centers_cart = np.asarray(centers_cart, dtype=float)
    wrapped = np.mod(centers_cart @ np.linalg.inv(cell), 1) @ cell
    starts, indices = csr(setup.tree, wrapped, setup.radius)
    n = len(wrapped)
    e = np.empty(n)
    gc = np.empty((n, 3))
    gu = np.empty((n, 3))
    arrays = [np.ascontiguousarray(wrapped), np.ascontiguousarray(axes, dtype=float),
              np.ascontiguousarray(setup.images), np.ascontiguousarray(setup.offsets),
              np.ascontiguousarray(setup.coeff12), np.ascontiguousarray(setup.coeff6),
              np.ascontiguousarray(setup.shift), np.ascontiguousarray(setup.floor2)]
    lib().oracle_pose_grad(n, *[_ptr(a) for a in arrays], setup.cutoff ** 2, _ptr(starts),
                           _ptr(indices), _ptr(e), _ptr(gc), _ptr(gu))
    return e, gc, gu


def string_pair(site, atom_frac, cell, eps_guest, sigma_guest, eps_frame, sigma_frame, cutoff):
    """native/string/source_code/triclinic_pair.h:14-68 restated (oracle C)."""
    cols = np.ascontiguousarray(np.asarray(cell, dtype=float).T)
    A, B, C = (np.ascontiguousarray(cols[i]) for i in range(3))
    return lib().oracle_string_pair(float(site[0]), float(site[1]), float(site[2]),
                                    float(atom_frac[0]), float(atom_frac[1]), float(atom_frac[2]),
                                    _ptr(A), _ptr(B), _ptr(C), eps_guest, sigma_guest, eps_frame,
                                    sigma_frame, cutoff)


def error_metrics(reference, candidate) -> dict:
    """tests/cof/reference_project/graph_vext/exact_neighbor_reuse_probe.py:73-83:
    accessible poses are E < 2000 K and |E| < 1e4 K; relative error floors |E| at 1 K.
    The extra floor-100 K figure is the mixed form's bound over every pose: below
    1e-5 it means at most 1e-3 K where |E| <= 100 K and 1e-5 relative above."""
    reference = np.asarray(reference, dtype=float)
    diff = np.abs(reference - np.asarray(candidate, dtype=float))
    acc = (reference < 2000.0) & (np.abs(reference) < 10000.0)
    rel = diff / np.maximum(np.abs(reference), 1.0)
    rel100 = diff / np.maximum(np.abs(reference), 100.0)
    return {"poses": int(diff.size), "accessible_poses": int(acc.sum()),
            "accessible_max_abs_error_K": float(diff[acc].max()) if acc.any() else None,
            "accessible_rms_error_K": float(np.sqrt((diff[acc] ** 2).mean())) if acc.any() else None,
            "all_max_relative_error_floor1K": float(rel.max()) if rel.size else 0.0,
            "all_max_relative_error_floor100K": float(rel100.max()) if rel100.size else 0.0,
            "all_max_abs_error_K": float(diff.max()) if diff.size else 0.0}
