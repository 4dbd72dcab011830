"""GPU evaluation behind the reference's native-kernel contract.

``NativeReusePotential(case, library, batch_size=512)`` has the interface of
``string_vext.native.NativeReusePotential`` (reference
src/string_vext/native.py:12-68): ``evaluate(centers, axes)`` takes fractional
centers and nonzero directors (normalized here) and returns FP64 energies
[center][axis]; ``query_seconds`` / ``compute_seconds`` accumulate the time
spent finding candidates and evaluating them, ``neighbor_queries`` /
``max_neighbors`` count the centers queried and the largest candidate set.

Two routes:
  * batched (the default when the library exports ``sv_plan_create``, as
    libsv_vext.so does): one framework plan and one pose set; the candidate
    lists are built on the GPU and one field launch covers every center;
  * drop-in (a library that exports only ``evaluate_reused_neighbors``, e.g.
    the reference CPU kernel, or ``batched=False``): centers go through the C
    ABI in blocks of ``batch_size``, each with a candidate CSR of the images
    within R = rc + max|o| of its centers, taken from a scipy sparse distance
    matrix between the block and the image tree.
"""
from __future__ import annotations

import ctypes
import time
from pathlib import Path

import numpy as np

from .neighbor import ReusePotential
from .physics import unit_directors

_ABI_ARGS = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double] + [ctypes.c_void_p] * 11


def candidate_csr(image_tree, points, radius):
    """CSR (starts int64 [n + 1], indices int32) of the images within `radius` of
    each point: the nonzeros of the sparse distance matrix points x images."""
    from scipy.spatial import cKDTree
    near = cKDTree(points).sparse_distance_matrix(image_tree, radius, output_type="coo_matrix")
    csr = near.tocsr()
    csr.sort_indices()
    return csr.indptr.astype(np.int64), csr.indices.astype(np.int32)


class NativeReusePotential(ReusePotential):
    def __init__(self, case: dict, library: Path, batch_size: int = 512, precision: str = "fp64",
                 batched: bool | None = None, near_energy_K: float | None = None):
        super().__init__(case)
        if batch_size < 1:
            raise ValueError("positive center batch required")
        self.batch_size = batch_size
        self.library = ctypes.CDLL(str(Path(library).resolve()))
        self.kernel = self.library.evaluate_reused_neighbors
        self.kernel.argtypes = _ABI_ARGS
        self.kernel.restype = ctypes.c_int
        can_batch = hasattr(self.library, "sv_plan_create")
        self.batched = can_batch if batched is None else bool(batched)
        if self.batched and not can_batch:
            raise ValueError("library has no batched entry points (sv_plan_create)")
        if precision != "fp64" and not self.batched:
            raise ValueError("mixed precision needs the batched entry points")
        self.precision = precision
        self.near_energy_K = near_energy_K
        self.query_seconds = 0.0
        self.compute_seconds = 0.0
        self._plan = None

    def plan(self):
        """The device plan of the batched route (built on first use)."""
        if self._plan is None:
            from .field import DEFAULT_NEAR_ENERGY_K, Plan
            near = DEFAULT_NEAR_ENERGY_K if self.near_energy_K is None else self.near_energy_K
            self._plan = Plan.from_case(self.case, near_energy_K=near)
        return self._plan

    def evaluate(self, centers: np.ndarray, axes: np.ndarray) -> np.ndarray:
        centers, axes = unit_directors(centers, axes,
                                       "finite fractional centers and nonzero directors required")
        route = self._evaluate_batched if self.batched else self._evaluate_blocks
        energy = route(np.mod(centers, 1), axes)
        if not np.isfinite(energy).all():
            raise ValueError("nonfinite native Vext energy")
        return energy

    def _evaluate_batched(self, wrapped, axes):
        tic = time.perf_counter()
        plan = self.plan()
        energy = plan.field(plan.poses(wrapped, axes), precision=self.precision)["energy"]
        self.neighbor_queries += len(wrapped)
        self.compute_seconds += time.perf_counter() - tic
        return energy

    def _evaluate_blocks(self, wrapped, axes):
        """One evaluate_reused_neighbors call per block of batch_size centers, with
        the image and coefficient tables passed on every call."""
        pot = self.potential
        cart = np.ascontiguousarray(wrapped @ pot.cell)
        tables = [np.ascontiguousarray(t) for t in (pot.images, self.offsets, self.coeff12,
                                                    self.coeff6, self.shift, self.floor2)]
        energy = np.empty((len(cart), len(axes)), dtype=np.float64)
        blocks = range(0, len(cart), self.batch_size)
        for lo, hi in zip(blocks, [*blocks[1:], len(cart)]):
            tic = time.perf_counter()
            starts, indices = candidate_csr(pot.tree, cart[lo:hi], self.radius)
            per_center = np.diff(starts)
            self.neighbor_queries += hi - lo
            self.max_neighbors = max(self.max_neighbors, int(per_center.max(initial=0)))
            self.query_seconds += time.perf_counter() - tic
            tic = time.perf_counter()
            args = [cart[lo:hi], axes, *tables, starts, indices, energy[lo:hi]]
            code = self.kernel(hi - lo, len(axes), len(pot.images), self.cutoff ** 2,
                               *(a.ctypes.data_as(ctypes.c_void_p) for a in args))
            self.compute_seconds += time.perf_counter() - tic
            if code:
                raise RuntimeError(f"native neighbor kernel failed with code {code}")
        return energy
