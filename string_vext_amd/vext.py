"""Exact triclinic orientation-resolved Vext field generation (``string-vext``).

Mirrors ``string_vext.vext`` (reference src/string_vext/vext.py:21-80): the
same NPZ channels, dtypes and layouts --
  orientation_K [K][n][n][n] float32, marginal_K [n][n][n] float32,
  site_K [2][n][n][n] float32, axes [K][3] float32, grid = asinh(marginal/1000),
  cell_matrix, atom_fractional, framework_sigma_A, framework_epsilon_K (FP64) --
the same CLI flags, exclusive output creation and provenance JSON.  With the
default kernel (libsv_vext.so) every channel is produced on the GPU in one
fused evaluation; ``--precision mixed`` selects the FP32-pair / FP64-accumulate
kernel.
"""
from __future__ import annotations

import argparse
import hashlib
import json
import os
from pathlib import Path

import numpy as np

from .input import parse_string_input
from .orientation import regular_pose_grid

DEFAULT_KERNEL = Path(__file__).resolve().parent / "libsv_vext.so"


def _has_batched(kernel: Path) -> bool:
    import ctypes
    return hasattr(ctypes.CDLL(str(Path(kernel).resolve())), "sv_plan_create")


def calculate_vext(case: dict, size: int = 60, orientations: int = 64,
                   kernel: Path = DEFAULT_KERNEL, precision: str = "fp64",
                   near_energy_K: float | None = None) -> dict[str, np.ndarray]:
    from .field import DEFAULT_NEAR_ENERGY_K, Plan
    from .physics import TriclinicSitePotential, marginal_energy, validate_case
    centers, axes = regular_pose_grid(size, orientations)
    validate_case(case)
    near = DEFAULT_NEAR_ENERGY_K if near_energy_K is None else near_energy_K
    if _has_batched(kernel):
        plan = Plan.from_case(case, near_energy_K=near)
        out = plan.field(plan.grid(size, axes), precision=precision, energy=False,
                         orientation=True, marginal=True, grid=True, site=True,
                         temperature=case["temperature"])
        orientation = out["orientation"].reshape(orientations, size, size, size)
        marginal = out["marginal"].reshape(size, size, size)
        site = out["site"].reshape(size, size, size)
        grid = out["grid"].reshape(size, size, size)
    else:
        from .native import NativeReusePotential
        energy = NativeReusePotential(case, kernel).evaluate(centers, axes)
        orientation = energy.T.reshape(orientations, size, size, size).astype(np.float32)
        marginal = marginal_energy(energy, case["temperature"]).reshape(size, size, size)
        marginal = marginal.astype(np.float32)
        site = TriclinicSitePotential(case).evaluate(
            centers @ case["cell"], case["guest_sigma"][0], case["guest_epsilon"][0])
        site = site.reshape(size, size, size).astype(np.float32)
        grid = np.arcsinh(marginal / 1000).astype(np.float32)
    result = {
        "orientation_K": orientation,
        "marginal_K": marginal,
        "site_K": np.stack((site, site)).astype(np.float32),
        "axes": axes.astype(np.float32),
        "cell_matrix": np.asarray(case["cell"], dtype=np.float64),
        "atom_fractional": np.asarray(case["frame_frac"], dtype=np.float64),
        "framework_sigma_A": np.asarray(case["frame_sigma"], dtype=np.float64),
        "framework_epsilon_K": np.asarray(case["frame_epsilon"], dtype=np.float64),
        "grid": grid,
    }
    if any(not np.isfinite(value).all() for value in result.values()):
        raise ValueError("nonfinite Vext output")
    return result


def main(argv=None):
    parser = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    parser.add_argument("--input", type=Path, required=True, help="native String input.dat")
    parser.add_argument("--output", type=Path, required=True, help="new NPZ destination")
    parser.add_argument("--size", type=int, default=60, help="grid points per axis")
    parser.add_argument("--orientations", type=int, default=64)
    parser.add_argument("--kernel", type=Path, default=DEFAULT_KERNEL)
    parser.add_argument("--precision", choices=("fp64", "mixed"), default="fp64")
    parser.add_argument("--near-energy", type=float, default=None,
                        help="mixed precision: pair terms that can exceed this (K) run in FP64")
    args = parser.parse_args(argv)
    record_path = args.output.with_suffix(".json")
    if args.output.exists() or record_path.exists():
        raise FileExistsError("Vext output or provenance record already exists")
    case = parse_string_input(args.input)
    fields = calculate_vext(case, args.size, args.orientations, args.kernel, args.precision,
                            args.near_energy)
    args.output.parent.mkdir(parents=True, exist_ok=True)
    temporary = args.output.with_suffix(".npz.tmp")
    try:
        with temporary.open("xb") as handle:
            np.savez_compressed(handle, **fields)
        os.link(temporary, args.output)
    finally:
        temporary.unlink(missing_ok=True)
    record = {"input": str(args.input.resolve()),
              "input_sha256": hashlib.sha256(args.input.read_bytes()).hexdigest(),
              "output_sha256": hashlib.sha256(args.output.read_bytes()).hexdigest(),
              "kernel_sha256": hashlib.sha256(args.kernel.read_bytes()).hexdigest(),
              "size": args.size, "orientations": args.orientations,
              "energy_mode": "exact_triclinic_LJ_neighbor_reuse",
              "precision": args.precision,
              "cell_lengths_A": case["cell_lengths"],
              "cell_angles_deg": case["cell_angles_deg"],
              "guest_sigma_A": case["guest_sigma"].tolist(),
              "guest_epsilon_K": case["guest_epsilon"].tolist()}
    with record_path.open("x") as handle:
        json.dump(record, handle, indent=2)
        handle.write("\n")
    print(json.dumps(record, sort_keys=True))


if __name__ == "__main__":
    main()
