"""Python handles over libsv_vext.so: framework plans, pose sets and batches.

    plan = Plan.from_case(case)                   # device tables of one framework+guest
    poses = plan.grid(60, axes)                   # 60^3 centers x len(axes) directors
    out = plan.field(poses, precision="mixed", temperature=298.0)

Host arrays in, host arrays out for ``Plan.field``; ``Batch`` keeps device
outputs resident for timing loops and many-field launches.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from .physics import body_offsets

# "fp64_rcp" (fields and batches only): FP64 with 1/d^2 from the hardware
# reciprocal and one Newton step instead of IEEE division.
PRECISIONS = {"fp64": 0, "mixed": 1, "fp64_rcp": 2}
POSE_PRECISIONS = ("fp64", "mixed")
DEFAULT_NEAR_ENERGY_K = 50.0

_P = ctypes.c_void_p


def _ptr(a):
    return None if a is None else a.ctypes.data_as(_P)


def _f64(a, shape=None):
    a = np.ascontiguousarray(a, dtype=np.float64)
    if shape is not None and a.shape != shape:
        raise ValueError(f"expected shape {shape}, got {a.shape}")
    return a


def precision_code(precision, allowed=tuple(PRECISIONS)) -> int:
    if precision not in allowed:
        raise ValueError(f"precision must be one of {sorted(allowed)}")
    return PRECISIONS[precision]


class DeviceBuffer:
    """Device allocation owned by libsv_vext.so (for callers without torch)."""

    def __init__(self, nbytes: int):
        self.lib = _lib.load()
        self.nbytes = int(nbytes)
        handle = ctypes.c_void_p()
        _lib.check(self.lib.sv_device_malloc(ctypes.byref(handle), self.nbytes))
        self.ptr = handle.value

    def to_host(self, dtype, shape):
        out = np.empty(shape, dtype=dtype)
        if out.nbytes > self.nbytes:
            raise ValueError("host array larger than the device buffer")
        _lib.check(self.lib.sv_memcpy_d2h(_ptr(out), self.ptr, out.nbytes))
        return out

    def free(self):
        if getattr(self, "ptr", None):
            self.lib.sv_device_free(self.ptr)
            self.ptr = None

    def __del__(self):
        self.free()


class Plan:
    """Device-resident framework + two-site guest + cutoff (sv_plan_create)."""

    def __init__(self, cell, frac, sigma, epsilon, guest_sigma, guest_epsilon, offsets, cutoff,
                 near_energy_K=DEFAULT_NEAR_ENERGY_K):
        self.lib = _lib.load()
        cell = _f64(cell, (3, 3))
        frac = _f64(frac)
        if frac.ndim != 2 or frac.shape[1] != 3:
            raise ValueError("fractional framework coordinates must be (n, 3)")
        sigma = _f64(sigma, (len(frac),))
        epsilon = _f64(epsilon, (len(frac),))
        offsets = _f64(offsets, (2,))
        handle = ctypes.c_void_p()
        _lib.check(self.lib.sv_plan_create(_ptr(cell), len(frac), _ptr(frac), _ptr(sigma),
                                           _ptr(epsilon), float(guest_sigma), float(guest_epsilon),
                                           _ptr(offsets), float(cutoff), float(near_energy_K),
                                           ctypes.byref(handle)))
        self.handle = handle.value
        self.cell = cell
        self.cutoff = float(cutoff)
        self.offsets = offsets

    @classmethod
    def from_case(cls, case, near_energy_K=DEFAULT_NEAR_ENERGY_K):
        """A plan for a parsed input case (string_vext_amd.input.parse_string_input)."""
        sig = np.asarray(case["guest_sigma"], dtype=float)
        eps = np.asarray(case["guest_epsilon"], dtype=float)
        if sig.shape != (2,) or eps.shape != (2,) or sig[0] != sig[1] or eps[0] != eps[1]:
            raise ValueError("exact neighbor reuse requires identical two-site LJ parameters")
        return cls(case["cell"], case["frame_frac"], case["frame_sigma"], case["frame_epsilon"],
                   sig[0], eps[0], body_offsets(case), case["cutoff"], near_energy_K)

    @property
    def n_types(self) -> int:
        n = ctypes.c_int()
        _lib.check(self.lib.sv_plan_info(self.handle, ctypes.byref(n), None))
        return n.value

    def poses(self, centers_frac, axes, tile_centers=64) -> "PoseSet":
        return PoseSet(self, centers_frac=centers_frac, axes=axes, tile_centers=tile_centers)

    def grid(self, size, axes, tile_centers=64) -> "PoseSet":
        return PoseSet(self, size=size, axes=axes, tile_centers=tile_centers)

    def field(self, poses: "PoseSet", precision="fp64", temperature=None, energy=True,
              orientation=False, marginal=False, grid=False, site=False) -> dict:
        """Evaluate every (center, axis) pose of `poses`; returns the requested arrays."""
        nc, nk = poses.n_centers, poses.n_axes
        out = {}
        if energy:
            out["energy"] = np.empty((nc, nk), dtype=np.float64)
        if orientation:
            out["orientation"] = np.empty((nk, nc), dtype=np.float32)
        if marginal or grid:
            out["marginal"] = np.empty(nc, dtype=np.float32)
        if grid:
            out["grid"] = np.empty(nc, dtype=np.float32)
        if site:
            out["site"] = np.empty(nc, dtype=np.float32)
        temp = float(temperature) if temperature is not None else 0.0
        _lib.check(self.lib.sv_field_run(self.handle, poses.handle, precision_code(precision), temp,
                                         _ptr(out.get("energy")), _ptr(out.get("orientation")),
                                         _ptr(out.get("marginal")), _ptr(out.get("grid")),
                                         _ptr(out.get("site"))))
        return out

    def census(self, poses: "PoseSet") -> int:
        """Exact number of in-cutoff (site, image) pairs over all poses."""
        n = ctypes.c_ulonglong()
        _lib.check(self.lib.sv_field_census(self.handle, poses.handle, ctypes.byref(n)))
        return int(n.value)

    def close(self):
        if getattr(self, "handle", None):
            self.lib.sv_plan_destroy(self.handle)
            self.handle = None

    def __del__(self):
        self.close()


class PoseSet:
    """Centers x unit directors on the device (sv_poses_create / sv_poses_grid)."""

    def __init__(self, plan: Plan, centers_frac=None, axes=None, size=None, tile_centers=64):
        self.lib = plan.lib
        self.plan = plan
        axes = _f64(axes)
        if axes.ndim != 2 or axes.shape[1] != 3 or len(axes) < 1:
            raise ValueError("axes must be (K, 3)")
        handle = ctypes.c_void_p()
        if size is not None:
            _lib.check(self.lib.sv_poses_grid_ex(plan.handle, int(size), len(axes), _ptr(axes),
                                                 int(tile_centers), ctypes.byref(handle)))
            self.n_centers = int(size) ** 3
        else:
            centers = _f64(centers_frac)
            if centers.ndim != 2 or centers.shape[1] != 3:
                raise ValueError("centers must be (n, 3) fractional coordinates")
            _lib.check(self.lib.sv_poses_create_ex(plan.handle, len(centers), _ptr(centers),
                                                   len(axes), _ptr(axes), int(tile_centers),
                                                   ctypes.byref(handle)))
            self.n_centers = len(centers)
        self.tile_centers = int(tile_centers)
        self.handle = handle.value
        self.n_axes = len(axes)

    def close(self):
        if getattr(self, "handle", None):
            self.lib.sv_poses_destroy(self.handle)
            self.handle = None

    def __del__(self):
        self.close()


class Batch:
    """Many (plan, poses) fields in one set of launches, outputs resident on the device.

    ``outputs`` selects the per-job device outputs among
    {"energy", "orientation", "marginal", "grid", "site"}.
    """

    def __init__(self, jobs, precision="mixed", temperatures=None,
                 outputs=("orientation", "marginal", "grid", "site"), external=None):
        """``external`` maps an output name to per-job device pointers owned by the
        caller (e.g. torch ``data_ptr()`` values); those outputs are not allocated here."""
        if not jobs:
            raise ValueError("at least one (plan, poses) job required")
        self.lib = _lib.load()
        self.jobs = list(jobs)
        n = len(self.jobs)
        external = external or {}
        self.buffers = {name: [] for name in ("energy", "orientation", "marginal", "grid", "site")}
        ptr_arrays = {}
        for name in self.buffers:
            ptrs = (ctypes.c_void_p * n)()
            if name in external:
                for j in range(n):
                    ptrs[j] = int(external[name][j])
                    self.buffers[name].append(None)
                ptr_arrays[name] = ptrs
                continue
            for j, (plan, poses) in enumerate(self.jobs):
                if name in outputs:
                    count = poses.n_centers * (poses.n_axes if name in ("energy", "orientation") else 1)
                    nbytes = count * (8 if name == "energy" else 4)
                    buf = DeviceBuffer(nbytes)
                    self.buffers[name].append(buf)
                    ptrs[j] = buf.ptr
                else:
                    self.buffers[name].append(None)
                    ptrs[j] = None
            ptr_arrays[name] = ptrs if name in outputs else None
        self.precision = precision
        plans = (ctypes.c_void_p * n)(*[p.handle for p, _ in self.jobs])
        poses = (ctypes.c_void_p * n)(*[s.handle for _, s in self.jobs])
        temps = np.ascontiguousarray(
            np.zeros(n) if temperatures is None else np.broadcast_to(temperatures, (n,)),
            dtype=np.float64)
        handle = ctypes.c_void_p()
        _lib.check(self.lib.sv_batch_create(n, plans, poses, precision_code(precision), _ptr(temps),
                                            ptr_arrays["energy"], ptr_arrays["orientation"],
                                            ptr_arrays["marginal"], ptr_arrays["grid"],
                                            ptr_arrays["site"], ctypes.byref(handle)))
        self.handle = handle.value

    def set_persistent(self, blocks_per_cu=8):
        """Launch the mixed field as one persistent grid (0 disables)."""
        _lib.check(self.lib.sv_batch_set_persistent(self.handle, int(blocks_per_cu)))

    def run(self, time_kernels=False, sync=True):
        _lib.check(self.lib.sv_batch_run(self.handle, int(bool(time_kernels)), int(bool(sync))))

    def sync(self):
        _lib.check(self.lib.sv_batch_sync(self.handle))

    def timings(self) -> dict:
        ms = np.zeros(7)
        _lib.check(self.lib.sv_batch_timings(self.handle, _ptr(ms)))
        return dict(zip(("list_count", "scan", "list_fill", "field", "marginal", "site", "run"),
                        ms.tolist()))

    def census(self) -> np.ndarray:
        out = np.zeros(len(self.jobs), dtype=np.uint64)
        _lib.check(self.lib.sv_batch_census(self.handle, _ptr(out)))
        return out

    def fetch(self, name, job=0) -> np.ndarray:
        plan, poses = self.jobs[job]
        buf = self.buffers[name][job]
        if buf is None:
            raise KeyError(f"output {name!r} was not requested")
        if name == "energy":
            return buf.to_host(np.float64, (poses.n_centers, poses.n_axes))
        if name == "orientation":
            return buf.to_host(np.float32, (poses.n_axes, poses.n_centers))
        return buf.to_host(np.float32, (poses.n_centers,))

    def close(self):
        if getattr(self, "handle", None):
            self.lib.sv_batch_destroy(self.handle)
            self.handle = None
        for bufs in getattr(self, "buffers", {}).values():
            for b in bufs:
                if b is not None:
                    b.free()

    def __del__(self):
        self.close()


class PoseGrid:
    """Device-resident independent poses (sv_pose_grid_*).

    A fixed grid of bins over the cell, sized for the two site positions of
    about ``n_expected`` poses (``tile_centers`` sites per bin, or 0 for the
    library's choice), whose candidate lists are built once.  ``eval``
    takes device pointers (e.g. torch ``data_ptr()``) of FP64 ``centers_cart``
    and ``axes`` arrays of shape (n, 3), bins the poses' sites on the GPU,
    evaluates every site and sums each pose's two, writing the requested
    device outputs; the poses never visit the host.
    """

    def __init__(self, plan: Plan, n_expected: int, tile_centers: int = 0):
        self.lib = plan.lib
        self.plan = plan
        handle = ctypes.c_void_p()
        _lib.check(self.lib.sv_pose_grid_create(plan.handle, int(n_expected), int(tile_centers),
                                                ctypes.byref(handle)))
        self.handle = handle.value
        # identifies the poses binned by the last eval (``token`` argument)
        self.binned_token = None

    def eval(self, n, centers, axes, precision="fp64", grad_out=None, energy=None,
             grad_center=None, grad_axis=None, token=None):
        """All arguments after ``n`` are device pointers (ints) or None."""
        self.binned_token = None
        _lib.check(self.lib.sv_pose_grid_eval(self.handle, int(n), centers, axes,
                                              precision_code(precision, POSE_PRECISIONS), grad_out, energy,
                                              grad_center, grad_axis))
        self.binned_token = token

    def eval_async(self, stream, n, centers, axes, precision="fp64", grad_out=None, energy=None,
                   grad_center=None, grad_axis=None, token=None):
        """Enqueue on ``stream`` (a raw hipStream_t handle, e.g. torch's
        ``current_stream().cuda_stream``; 0 = default stream) without waiting.
        Invalid input and nonfinite results are reported by ``check``.
        ``token`` (any object) is kept as ``binned_token`` to name these poses."""
        self.binned_token = None
        _lib.check(self.lib.sv_pose_grid_eval_async(self.handle, stream or None, int(n), centers,
                                                    axes, precision_code(precision, POSE_PRECISIONS), grad_out,
                                                    energy, grad_center, grad_axis))
        self.binned_token = token

    def reeval_async(self, stream, precision="fp64", grad_out=None, energy=None,
                     grad_center=None, grad_axis=None):
        """Evaluate the poses of the last eval again (new outputs) from the
        grid's binned copies, without binning them again."""
        _lib.check(self.lib.sv_pose_grid_reeval_async(self.handle, stream or None,
                                                      precision_code(precision, POSE_PRECISIONS), grad_out, energy,
                                                      grad_center, grad_axis))

    def check(self):
        """Wait for the last enqueued evaluation; raise on flagged input/results."""
        _lib.check(self.lib.sv_pose_grid_check(self.handle))

    def census(self) -> int:
        """Exact in-cutoff (site, image) pairs of the last binned poses."""
        n = ctypes.c_ulonglong()
        _lib.check(self.lib.sv_pose_grid_census(self.handle, ctypes.byref(n)))
        return int(n.value)

    def set_parts(self, parts):
        """Blocks per run of up to 64 sites (1..8): more spread a call's work over
        more blocks at a fixed cost per block."""
        _lib.check(self.lib.sv_pose_grid_set_parts(self.handle, int(parts)))

    def set_timing(self, on=True):
        """Record HIP events around the phases of every later evaluation."""
        _lib.check(self.lib.sv_pose_grid_set_timing(self.handle, int(bool(on))))

    def timings(self) -> dict:
        """Milliseconds of the last timed evaluation's phases (waits for it):
        binning (0 for a re-evaluation), site kernel, pose sums."""
        ms = np.zeros(3)
        _lib.check(self.lib.sv_pose_grid_timings(self.handle, _ptr(ms)))
        return dict(zip(("binning", "site_kernel", "combine"), ms.tolist()))

    def info(self):
        tiles, cap = ctypes.c_int(), ctypes.c_int()
        _lib.check(self.lib.sv_pose_grid_info(self.handle, ctypes.byref(tiles), ctypes.byref(cap)))
        return {"tiles": tiles.value, "capacity": cap.value}

    def close(self):
        if getattr(self, "handle", None):
            self.lib.sv_pose_grid_destroy(self.handle)
            self.handle = None

    def __del__(self):
        self.close()


class PoseList:
    """Independent poses (center_i, director_i) on the device (sv_pose_list_create).

    ``centers_cart`` are Cartesian (Angstrom); directors are used as given.
    ``eval`` writes device arrays: pass raw device pointers (e.g. torch
    ``data_ptr()``) or leave them ``None`` to skip an output.
    """

    def __init__(self, plan: Plan, centers_cart, axes):
        self.lib = plan.lib
        self.plan = plan
        centers = _f64(centers_cart)
        axes = _f64(axes)
        if centers.ndim != 2 or centers.shape[1] != 3 or axes.shape != centers.shape:
            raise ValueError("centers and axes must both be (n, 3)")
        handle = ctypes.c_void_p()
        _lib.check(self.lib.sv_pose_list_create(plan.handle, len(centers), _ptr(centers),
                                                _ptr(axes), ctypes.byref(handle)))
        self.handle = handle.value
        self.n = len(centers)

    def eval(self, precision="fp64", grad_out=None, energy=None, grad_center=None, grad_axis=None):
        """Device pointers (ints) for every array; returns nothing."""
        _lib.check(self.lib.sv_pose_list_eval(self.handle, precision_code(precision, POSE_PRECISIONS), grad_out,
                                              energy, grad_center, grad_axis))

    def energy_host(self, precision="fp64", grad=False):
        """Convenience: energies (and gradients) copied back to numpy."""
        bufs = [DeviceBuffer(8 * self.n)]
        if grad:
            bufs += [DeviceBuffer(24 * self.n), DeviceBuffer(24 * self.n)]
        self.eval(precision, None, bufs[0].ptr, *(b.ptr for b in bufs[1:]))
        out = [bufs[0].to_host(np.float64, (self.n,))]
        if grad:
            out += [b.to_host(np.float64, (self.n, 3)) for b in bufs[1:]]
        return out if grad else out[0]

    def close(self):
        if getattr(self, "handle", None):
            self.lib.sv_poses_destroy(self.handle)
            self.handle = None

    def __del__(self):
        self.close()
