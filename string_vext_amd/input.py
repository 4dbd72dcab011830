"""Parser of the native String ``input.dat`` format for neutral two-site guests.

Mirrors ``string_vext.input.parse_string_input`` (reference
src/string_vext/input.py:11-57): values follow header lines located by prefix;
framework rows are ``ID sigma epsilon charge mass fa fb fc name``, guest rows
``x y z epsilon sigma charge mass``; charged or malformed inputs are rejected;
the guest is re-expressed as a canonical body on the x axis about its centre of
mass.
"""
from __future__ import annotations

from pathlib import Path

import numpy as np

from .cell import cell_matrix


def parse_string_input(path) -> dict:
    lines = [line.strip() for line in Path(path).read_text().splitlines()]

    def after(prefix):
        for index, line in enumerate(lines):
            if line.startswith(prefix):
                return index + 1
        raise ValueError(f"input is missing the {prefix!r} header")

    lengths = [float(v) for v in lines[after("La Lb Lc")].split()[:3]]
    angles = [float(v) for v in lines[after("Alpha Beta Gamma")].split()[:3]]
    cell = cell_matrix(lengths, angles)
    conditions = [float(v) for v in lines[after("cutoff(A)")].split()]
    g = after("Number of sites")
    n_guest = int(lines[g])
    guest = np.array([[float(v) for v in row.split()[:7]] for row in lines[g + 2:g + 2 + n_guest]])
    f = after("Number of atoms")
    n_frame = int(lines[f])
    frame = np.array([[float(v) for v in row.split()[:8]] for row in lines[f + 2:f + 2 + n_frame]])
    if (guest.shape != (2, 7) or frame.shape != (n_frame, 8) or len(conditions) < 4
            or not np.isfinite(guest).all() or not np.isfinite(frame).all()
            or not np.isfinite(conditions).all()
            or not np.array_equal(frame[:, 0], np.arange(1, n_frame + 1))
            or (guest[:, 5] != 0).any() or (frame[:, 3] != 0).any()
            or (guest[:, 3:5] <= 0).any() or (guest[:, 6] <= 0).any()
            or (frame[:, 1:3] <= 0).any() or conditions[0] <= 0 or conditions[3] <= 0):
        raise ValueError("invalid or charged two-site String input")
    mass = guest[:, 6]
    if not np.isclose(mass.sum(), conditions[2], rtol=1e-7):
        raise ValueError("guest site masses differ from declared total mass")
    body = guest[:, :3]
    separation = body[1] - body[0]
    bond = np.linalg.norm(separation)
    if bond <= 0:
        raise ValueError("two distinct guest sites required")
    director = separation / bond
    com = (body * mass[:, None]).sum(0) / conditions[2]
    canonical = np.zeros((2, 3))
    canonical[:, 0] = (body - com) @ director
    return {"cell": cell, "frame_frac": np.mod(frame[:, 5:8], 1),
            "frame_sigma": frame[:, 1], "frame_epsilon": frame[:, 2],
            "guest_xyz": canonical, "guest_sigma": guest[:, 4], "guest_epsilon": guest[:, 3],
            "guest_mass": mass, "total_mass": float(conditions[2]),
            "temperature": float(conditions[3]), "cutoff": float(conditions[0]),
            "cell_lengths": lengths, "cell_angles_deg": angles, "bond_length": float(bond)}


def write_string_input(case: dict, path) -> None:
    """Write a case in the input.dat layout that parse_string_input reads back."""
    lengths = case.get("cell_lengths")
    angles = case.get("cell_angles_deg")
    if lengths is None or angles is None:
        raise ValueError("case needs cell_lengths and cell_angles_deg")
    xyz = np.asarray(case["guest_xyz"], dtype=float)
    rows = ["Nmaxa Nmaxb Nmaxc:", "20 20 20", "La Lb Lc dL",
            f"{float(lengths[0])!r} {float(lengths[1])!r} {float(lengths[2])!r} 0.5", "Alpha Beta Gamma",
            f"{float(angles[0])!r} {float(angles[1])!r} {float(angles[2])!r}",
            "cutoff(A) FH_signal Mass(g/mol) Tempearture(K) Running_steps",
            f"{float(case['cutoff'])!r} 1 {float(case['total_mass'])!r} {float(case['temperature'])!r} 100",
            "---------String Calculation Settings---------", "Direction", "1",
            "#_of_points delta_frac delta_angle_degree", "401 0.0001 1.0", "convergence_setting",
            "default", "------------------Adsorbate------------------", "Number of sites", "2",
            "x y z Epsilon Sigma Charge Mass"]
    for s in range(2):
        rows.append(f"{float(xyz[s, 0])!r} 0 0 {float(case['guest_epsilon'][s])!r} "
                    f"{float(case['guest_sigma'][s])!r} 0 {float(case['guest_mass'][s])!r}")
    rows += ["------------------Adsorbent-----------------", "Number of atoms",
             str(len(case["frame_frac"])), "ID diameter Epsilon Charge mass frac_x frac_y frac_z atom_name"]
    for i, (f, s, e) in enumerate(zip(case["frame_frac"], case["frame_sigma"], case["frame_epsilon"])):
        rows.append(f"{i + 1} {float(s)!r} {float(e)!r} 0 12.0 {float(f[0])!r} {float(f[1])!r} "
                    f"{float(f[2])!r} C")
    Path(path).write_text("\n".join(rows) + "\n")
