"""Exact triclinic orientation-resolved Vext field generation (``string-vext``).

Contract of ``string_vext.vext`` (reference src/string_vext/vext.py:21-80):
``calculate_vext(case, size, orientations, kernel)`` returns the NPZ channels

  orientation_K [K][n][n][n] float32   pose energies, orientation-major
  marginal_K    [n][n][n]    float32   -T ln(mean_k exp(-E_k / T))
  site_K        [2][n][n][n] float32   single-site field, once per site
  axes          [K][3]       float32   Fibonacci directors
  grid          [n][n][n]    float32   asinh(marginal / 1000)
  cell_matrix, atom_fractional, framework_sigma_A, framework_epsilon_K  (FP64)

and ``main`` is the CLI (--input --output --size --orientations --kernel),
which refuses to overwrite, publishes the NPZ atomically under its final
name, and writes a provenance JSON beside it.  With the default kernel
(libsv_vext.so) every channel comes from one fused GPU evaluation; with a
library that only exports ``evaluate_reused_neighbors`` the energies go
through that C ABI in 512-center blocks.  Added here: ``--precision mixed``
(FP32 pair arithmetic, FP64 accumulation), ``--precision fp64_rcp`` (FP64
with the hardware reciprocal) and ``--near-energy``.
"""
from __future__ import annotations

import argparse
import ctypes
import hashlib
import json
import os
import tempfile
from pathlib import Path

import numpy as np

from .input import parse_string_input
from .orientation import regular_pose_grid

DEFAULT_KERNEL = Path(__file__).resolve().parent / "libsv_vext.so"
ENERGY_MODE = "exact_triclinic_LJ_neighbor_reuse"


def _exports_plans(kernel: Path) -> bool:
    return hasattr(ctypes.CDLL(str(Path(kernel).resolve())), "sv_plan_create")


def _gpu_channels(case, size, axes, precision, near_energy_K):
    """orientation [K][n^3], marginal, grid, site [n^3] from one fused launch."""
    from .field import Plan
    plan = Plan.from_case(case, near_energy_K=near_energy_K)
    out = plan.field(plan.grid(size, axes), precision=precision, energy=False, orientation=True,
                     marginal=True, grid=True, site=True, temperature=case["temperature"])
    return out["orientation"], out["marginal"], out["grid"], out["site"]


def _abi_channels(case, centers, axes, kernel):
    """The same channels with the energies from the drop-in C ABI."""
    from .native import NativeReusePotential
    from .physics import TriclinicSitePotential, marginal_energy
    energy = NativeReusePotential(case, kernel).evaluate(centers, axes)
    marginal = marginal_energy(energy, case["temperature"]).astype(np.float32)
    site = TriclinicSitePotential(case).evaluate(centers @ case["cell"], case["guest_sigma"][0],
                                                 case["guest_epsilon"][0])
    return (energy.T.astype(np.float32), marginal, np.arcsinh(marginal / 1000).astype(np.float32),
            site.astype(np.float32))


def calculate_vext(case: dict, size: int = 60, orientations: int = 64,
                   kernel: Path = DEFAULT_KERNEL, precision: str = "fp64",
                   near_energy_K: float | None = None) -> dict[str, np.ndarray]:
    from .field import DEFAULT_NEAR_ENERGY_K
    from .physics import validate_case
    validate_case(case)
    centers, axes = regular_pose_grid(size, orientations)
    if _exports_plans(kernel):
        near = DEFAULT_NEAR_ENERGY_K if near_energy_K is None else near_energy_K
        orientation, marginal, grid, site = _gpu_channels(case, size, axes, precision, near)
    else:
        if precision != "fp64":
            raise ValueError(f"precision {precision!r} needs a kernel library with the batched "
                             "entry points (sv_plan_create); this one only evaluates FP64")
        orientation, marginal, grid, site = _abi_channels(case, centers, axes, kernel)
    volume = (size, size, size)
    fields = {
        "orientation_K": orientation.reshape(orientations, *volume),
        "marginal_K": marginal.reshape(volume),
        "site_K": np.broadcast_to(site.reshape(volume), (2, *volume)).astype(np.float32),
        "axes": axes.astype(np.float32),
        "grid": grid.reshape(volume),
    }
    for name, key in (("cell_matrix", "cell"), ("atom_fractional", "frame_frac"),
                      ("framework_sigma_A", "frame_sigma"), ("framework_epsilon_K", "frame_epsilon")):
        fields[name] = np.asarray(case[key], dtype=np.float64)
    if not all(np.isfinite(v).all() for v in fields.values()):
        raise ValueError("nonfinite Vext output")
    return fields


def _sha256(path: Path) -> str:
    digest = hashlib.sha256()
    with open(path, "rb") as handle:
        for block in iter(lambda: handle.read(1 << 20), b""):
            digest.update(block)
    return digest.hexdigest()


def _publish_npz(destination: Path, fields: dict) -> str:
    """Write the NPZ under a unique temporary name in the destination's folder
    and hard-link it to its final name: readers never see a partial file, and
    an existing destination makes the link (and the call) fail.  The archive
    is the format of numpy.savez_compressed, deflated in parallel chunks
    (string_vext_amd.npz).  Returns the SHA-256 of the file, hashed from the
    bytes as they are written."""
    from .npz import write_npz
    destination.parent.mkdir(parents=True, exist_ok=True)
    fd, staging = tempfile.mkstemp(prefix=destination.name + ".", suffix=".partial",
                                   dir=destination.parent)
    digest = hashlib.sha256()
    try:
        with os.fdopen(fd, "wb") as handle:
            write_npz(handle, fields, digest=digest)
        os.link(staging, destination)
    finally:
        os.unlink(staging)
    return digest.hexdigest()


def _provenance(args, case, output_sha256: str | None = None) -> dict:
    sources = {"input": args.input, "output": args.output, "kernel": args.kernel}
    record = {"input": str(args.input.resolve())}
    record.update({f"{name}_sha256": output_sha256 if name == "output" and output_sha256
                   else _sha256(path) for name, path in sources.items()})
    record.update(size=args.size, orientations=args.orientations, energy_mode=ENERGY_MODE,
                  precision=args.precision, cell_lengths_A=case["cell_lengths"],
                  cell_angles_deg=case["cell_angles_deg"],
                  guest_sigma_A=np.asarray(case["guest_sigma"]).tolist(),
                  guest_epsilon_K=np.asarray(case["guest_epsilon"]).tolist())
    return record


def _arguments(argv):
    cli = argparse.ArgumentParser(prog="string-vext", description=__doc__.splitlines()[0])
    cli.add_argument("--input", type=Path, required=True,
                     help="String input.dat (framework, guest, cell, cutoff, temperature)")
    cli.add_argument("--output", type=Path, required=True,
                     help="NPZ to create; it and its .json record must not exist yet")
    cli.add_argument("--size", type=int, default=60, help="centers per lattice direction")
    cli.add_argument("--orientations", type=int, default=64, help="Fibonacci directors")
    cli.add_argument("--kernel", type=Path, default=DEFAULT_KERNEL,
                     help="shared library exporting evaluate_reused_neighbors")
    cli.add_argument("--precision", choices=("fp64", "fp64_rcp", "mixed"), default="fp64",
                     help="fp64: the reference's arithmetic; fp64_rcp: FP64 with the hardware "
                          "reciprocal and one Newton step; mixed: FP32 pair terms, FP64 sums")
    cli.add_argument("--near-energy", type=float, default=None,
                     help="mixed precision: pair terms that can exceed this (K) run in FP64")
    return cli.parse_args(argv)


def main(argv=None):
    args = _arguments(argv)
    record_path = args.output.with_suffix(".json")
    if any(p.exists() for p in (args.output, record_path)):
        raise FileExistsError("Vext output or provenance record already exists")
    case = parse_string_input(args.input)
    output_sha256 = _publish_npz(args.output, calculate_vext(
        case, args.size, args.orientations, args.kernel, args.precision, args.near_energy))
    record = _provenance(args, case, output_sha256)
    with record_path.open("x") as handle:
        handle.write(json.dumps(record, indent=2) + "\n")
    print(json.dumps(record, sort_keys=True))


if __name__ == "__main__":
    main()
