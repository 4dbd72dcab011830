"""Loader of tests/golden/golden_vext.npz (written by tests/golden/make_golden.py)."""
from pathlib import Path

import numpy as np

GOLDEN = Path(__file__).resolve().parent / "golden" / "golden_vext.npz"
NAMES = ("toy", "c2h4", "c2h6")


def load(name):
    """(case, centers_frac, axes, reference_energy) of one golden case."""
    with np.load(GOLDEN, allow_pickle=False) as z:
        get = lambda key: z[f"{name}/{key}"]  # noqa: E731
        total_mass, temperature, cutoff = get("scalars")
        case = {key: get(key) for key in ("cell", "frame_frac", "frame_sigma", "frame_epsilon",
                                          "guest_xyz", "guest_sigma", "guest_epsilon",
                                          "guest_mass")}
        case.update(total_mass=float(total_mass), temperature=float(temperature),
                    cutoff=float(cutoff), cell_lengths=get("cell_lengths").tolist(),
                    cell_angles_deg=get("cell_angles_deg").tolist())
        return case, get("centers"), get("axes"), get("energy")
