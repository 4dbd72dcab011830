"""Reader and writer of the native String ``input.dat`` format for neutral
two-site guests.

The format (reference src/string_vext/input.py:11-57, read positionally by
the String solver, native/string/source_code/RC_main.cu:282-387) is a list of
header lines, each followed by its values:

    La Lb Lc ...           a b c (A)
    Alpha Beta Gamma       alpha beta gamma (deg)
    cutoff(A) ...          cutoff, FH signal, guest mass (g/mol), temperature (K), ...
    Number of sites        n, then a column header, then n rows  x y z eps sigma q m
    Number of atoms        n, then a column header, then n rows  ID sigma eps q m fa fb fc name

``parse_string_input`` accepts a neutral guest of exactly two sites and a
neutral framework with IDs 1..n, positive LJ parameters, masses, cutoff and
temperature, and a total mass equal to the site masses (relative 1e-7).  The
guest is returned as a canonical body on the x axis about its centre of mass.
"""
from __future__ import annotations

from pathlib import Path

import numpy as np

from .cell import cell_matrix

_INVALID = "invalid or charged two-site String input"

# guest row columns: x y z eps sigma q m
_G_XYZ, _G_EPS, _G_SIG, _G_Q, _G_MASS = slice(0, 3), 3, 4, 5, 6
# framework row columns: ID sigma eps q m fa fb fc
_F_ID, _F_SIG, _F_EPS, _F_Q, _F_FRAC = 0, 1, 2, 3, slice(5, 8)
# condition columns: cutoff FH mass temperature
_C_CUT, _C_MASS, _C_TEMP = 0, 2, 3


class _Sections:
    """The stripped lines of an input file, addressed by header prefix."""

    def __init__(self, text: str):
        self.lines = [line.strip() for line in text.splitlines()]

    def value_line(self, header: str) -> int:
        """Index of the line after the first line that starts with `header`."""
        hit = next((i for i, line in enumerate(self.lines) if line.startswith(header)), None)
        if hit is None:
            raise ValueError(f"input is missing the {header!r} header")
        return hit + 1

    def floats(self, header: str, count: int | None = None) -> list[float]:
        fields = self.lines[self.value_line(header)].split()
        return [float(v) for v in (fields if count is None else fields[:count])]

    def table(self, header: str, width: int) -> np.ndarray:
        """The rows under a count line: count, one column-title line, then rows
        (a file that ends early gives fewer rows)."""
        at = self.value_line(header)
        count = int(self.lines[at])
        body = self.lines[at + 2:at + 2 + count]
        rows = [[float(v) for v in row.split()[:width]] for row in body]
        return np.array(rows, dtype=float).reshape(len(rows), -1) if rows else np.empty((0, width))


def _well_formed(guest, frame, conditions, n_frame) -> bool:
    if guest.shape != (2, 7) or frame.shape != (n_frame, 8) or len(conditions) < 4:
        return False
    if not (np.isfinite(guest).all() and np.isfinite(frame).all() and np.isfinite(conditions).all()):
        return False
    ids_in_order = np.array_equal(frame[:, _F_ID], np.arange(1, n_frame + 1))
    neutral = not guest[:, _G_Q].any() and not frame[:, _F_Q].any()
    positive = ((guest[:, [_G_EPS, _G_SIG, _G_MASS]] > 0).all()
                and (frame[:, [_F_SIG, _F_EPS]] > 0).all()
                and conditions[_C_CUT] > 0 and conditions[_C_TEMP] > 0)
    return ids_in_order and neutral and positive


def _canonical_body(xyz, mass, total_mass):
    """Site x coordinates of the body laid on the x axis about its centre of mass,
    and the bond length."""
    bond_vector = xyz[1] - xyz[0]
    bond = float(np.linalg.norm(bond_vector))
    if bond <= 0:
        raise ValueError("two distinct guest sites required")
    com = (mass @ xyz) / total_mass
    body = np.zeros((2, 3))
    body[:, 0] = (xyz - com) @ (bond_vector / bond)
    return body, bond


def parse_string_input(path) -> dict:
    sec = _Sections(Path(path).read_text())
    lengths = sec.floats("La Lb Lc", 3)
    angles = sec.floats("Alpha Beta Gamma", 3)
    cell = cell_matrix(lengths, angles)
    conditions = sec.floats("cutoff(A)")
    n_frame = int(sec.lines[sec.value_line("Number of atoms")])
    guest = sec.table("Number of sites", 7)
    frame = sec.table("Number of atoms", 8)
    if not _well_formed(guest, frame, conditions, n_frame):
        raise ValueError(_INVALID)
    mass = guest[:, _G_MASS]
    total = conditions[_C_MASS]
    if not np.isclose(mass.sum(), total, rtol=1e-7):
        raise ValueError("guest site masses differ from declared total mass")
    body, bond = _canonical_body(guest[:, _G_XYZ], mass, total)
    return {"cell": cell, "frame_frac": np.mod(frame[:, _F_FRAC], 1),
            "frame_sigma": frame[:, _F_SIG], "frame_epsilon": frame[:, _F_EPS],
            "guest_xyz": body, "guest_sigma": guest[:, _G_SIG], "guest_epsilon": guest[:, _G_EPS],
            "guest_mass": mass, "total_mass": float(total),
            "temperature": float(conditions[_C_TEMP]), "cutoff": float(conditions[_C_CUT]),
            "cell_lengths": lengths, "cell_angles_deg": angles, "bond_length": bond}


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
