"""Writes tests/golden/golden_vext.npz: inputs and the reference kernel's outputs.

Run in a container that holds the reference tree (after `make -C oracle`):

    python tests/golden/make_golden.py

The energies come from oracle/_ref/kernel.so -- the reference's own
native/vext/kernel.cpp compiled from its source -- driven by the oracle's
restatement of the reference's host preparation (images, coefficients and
cKDTree CSR of src/string_vext/native.py:26-68).  Inputs:
  toy   examples/toy_triclinic/input.dat (the same case as
        tests/unit/test_public_api.py:13-39), 3^3 grid x 4 directions;
  c2h4, c2h6  tests/cof/examples/{c2h4,c2h6}/input.dat (69-atom nonorthogonal
        frameworks, c = 5.926 A < rc), the boundary centers of
        test_exact_neighbor_reuse_native.py:23-26 plus a 3^3 grid, with the
        axes x, y, z, (1,1,1) of the same test plus 8 Fibonacci directions.
"""
from __future__ import annotations

import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

from oracle import vext_oracle as O  # noqa: E402

REF = Path("/root/reference")
OUT = Path(__file__).resolve().parent / "golden_vext.npz"
KEYS = ("cell", "frame_frac", "frame_sigma", "frame_epsilon", "guest_xyz", "guest_sigma",
        "guest_epsilon", "guest_mass")


def record(store, name, case, centers, axes):
    energy = O.reuse_energy(case, centers, axes, kernel="reference")
    for key in KEYS:
        store[f"{name}/{key}"] = np.asarray(case[key], dtype=np.float64)
    store[f"{name}/scalars"] = np.array([case["total_mass"], case["temperature"], case["cutoff"]])
    store[f"{name}/cell_lengths"] = np.asarray(case["cell_lengths"], dtype=np.float64)
    store[f"{name}/cell_angles_deg"] = np.asarray(case["cell_angles_deg"], dtype=np.float64)
    store[f"{name}/centers"] = np.asarray(centers, dtype=np.float64)
    store[f"{name}/axes"] = np.asarray(axes, dtype=np.float64)
    store[f"{name}/energy"] = energy


def main():
    if O.reference_kernel() is None:
        raise SystemExit("oracle/_ref/kernel.so missing: run `make -C oracle` with the reference present")
    store = {}
    toy = O.parse_string_input(REF / "examples/toy_triclinic/input.dat")
    centers, axes = O.regular_pose_grid(3, 4)
    record(store, "toy", toy, centers, axes)
    edge = np.array([[0.001, 0.999, 0.5], [0.999, 0.001, 0.5], [0.25, 0.5, 0.75], [0.5, 0.5, 0.5]])
    grid, fib = O.regular_pose_grid(3, 8)
    axes = np.concatenate([np.array([[1.0, 0, 0], [0, 1.0, 0], [0, 0, 1.0], [1.0, 1.0, 1.0]]), fib])
    for gas in ("c2h4", "c2h6"):
        case = O.parse_string_input(REF / f"tests/cof/examples/{gas}/input.dat")
        record(store, gas, case, np.concatenate([edge, grid]), axes)
    np.savez_compressed(OUT, **store)
    print(OUT, OUT.stat().st_size, "bytes")


if __name__ == "__main__":
    main()
