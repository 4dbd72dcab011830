"""GPU evaluation behind the reference's native-kernel contract.

Mirrors ``string_vext.native.NativeReusePotential`` (reference
src/string_vext/native.py:12-68): ``evaluate(centers, axes)`` returns FP64
energies [center][axis] for fractional centers and (normalized) directors.

Two routes:
  * the batched route (default with libsv_vext.so): one plan + pose set, the
    candidate lists built on the device, one field launch for all centers;
  * the drop-in route (any library exporting only evaluate_reused_neighbors,
    e.g. the reference CPU kernel, or ``batched=False``): the reference's
    loop of 512-center batches with a cKDTree CSR, one C-ABI call per batch.
"""
from __future__ import annotations

import ctypes
import time
from pathlib import Path

import numpy as np

from .neighbor import ReusePotential


class NativeReusePotential(ReusePotential):
    def __init__(self, case: dict, library: Path, batch_size: int = 512, precision: str = "fp64",
                 batched: bool | None = None, near_energy_K: float | None = None):
        super().__init__(case)
        if batch_size < 1:
            raise ValueError("positive center batch required")
        self.batch_size = batch_size
        self.library = ctypes.CDLL(str(Path(library).resolve()))
        self.kernel = self.library.evaluate_reused_neighbors
        self.kernel.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                ctypes.c_double] + [ctypes.c_void_p] * 11
        self.kernel.restype = ctypes.c_int
        has_batched = hasattr(self.library, "sv_plan_create")
        self.batched = has_batched if batched is None else bool(batched)
        if self.batched and not has_batched:
            raise ValueError("library has no batched entry points (sv_plan_create)")
        if precision != "fp64" and not self.batched:
            raise ValueError("mixed precision needs the batched entry points")
        self.precision = precision
        self.near_energy_K = near_energy_K
        self.query_seconds = 0.0
        self.compute_seconds = 0.0
        self._plan = None

    def plan(self):
        if self._plan is None:
            from .field import DEFAULT_NEAR_ENERGY_K, Plan
            near = DEFAULT_NEAR_ENERGY_K if self.near_energy_K is None else self.near_energy_K
            self._plan = Plan.from_case(self.case, near_energy_K=near)
        return self._plan

    def evaluate(self, centers: np.ndarray, axes: np.ndarray) -> np.ndarray:
        centers, axes = np.asarray(centers, dtype=float), np.asarray(axes, dtype=float)
        if (centers.ndim != 2 or centers.shape[1] != 3
                or axes.ndim != 2 or axes.shape[1] != 3
                or not np.isfinite(centers).all() or not np.isfinite(axes).all()
                or (np.linalg.norm(axes, axis=1) <= 0).any()):
            raise ValueError("finite fractional centers and nonzero directors required")
        axes = np.ascontiguousarray(axes / np.linalg.norm(axes, axis=1, keepdims=True))
        if self.batched:
            stamp = time.perf_counter()
            plan = self.plan()
            output = plan.field(plan.poses(np.mod(centers, 1), axes),
                                precision=self.precision)["energy"]
            self.neighbor_queries += len(centers)
            self.compute_seconds += time.perf_counter() - stamp
        else:
            output = self._evaluate_csr(centers, axes)
        if not np.isfinite(output).all():
            raise ValueError("nonfinite native Vext energy")
        return output

    def _evaluate_csr(self, centers, axes):
        cart = np.ascontiguousarray(np.mod(centers, 1) @ self.potential.cell)
        buffers = tuple(np.ascontiguousarray(a) for a in (
            self.potential.images, self.offsets, self.coeff12, self.coeff6, self.shift,
            self.floor2))
        output = np.empty((len(centers), len(axes)), dtype=np.float64)
        for start in range(0, len(cart), self.batch_size):
            end = min(start + self.batch_size, len(cart))
            stamp = time.perf_counter()
            groups = self.potential.tree.query_ball_point(cart[start:end], self.radius, workers=1)
            lengths = np.fromiter((len(g) for g in groups), dtype=np.int64, count=len(groups))
            starts = np.empty(len(groups) + 1, dtype=np.int64)
            starts[0] = 0
            starts[1:] = np.cumsum(lengths)
            indices = (np.concatenate([np.asarray(g, dtype=np.int32) for g in groups])
                       if starts[-1] else np.empty(0, np.int32))
            self.neighbor_queries += len(groups)
            self.max_neighbors = max(self.max_neighbors, int(lengths.max(initial=0)))
            self.query_seconds += time.perf_counter() - stamp
            stamp = time.perf_counter()
            arrays = (cart[start:end], axes, *buffers, starts, indices, output[start:end])
            result = self.kernel(len(groups), len(axes), len(self.potential.images),
                                 self.cutoff ** 2,
                                 *(a.ctypes.data_as(ctypes.c_void_p) for a in arrays))
            if result:
                raise RuntimeError(f"native neighbor kernel failed with code {result}")
            self.compute_seconds += time.perf_counter() - stamp
        return output
