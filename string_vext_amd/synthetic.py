"""Seeded synthetic frameworks and guests for the benchmark workloads.

Each generator returns a ``case`` dict with the keys of
``string_vext_amd.input.parse_string_input``.

  small_field_case   25 x 25 x 12 A, gamma 120, 300 atoms of 3 types
                     (eps 28/35/8 K, sigma 3.40/3.26/2.85 A), rejection to a
                     1.2 A minimum image distance; 20^3 x 64 field
  full_field_case    37 x 37 x 14 A, gamma 120, 1200 atoms of 4 types
                     (eps ~ U(8, 60) K, sigma ~ U(2.5, 3.8) A), 1.2 A rejection;
                     60^3 x 64 field
  ragged_cases       frameworks with N ~ logU(200, 4000), lengths ~ U(10, 60) A,
                     angles ~ U(60, 120) (positive volume), uniform fractional
                     coordinates, 4 LJ types as above, x guests with
                     eps ~ U(60, 250) K, sigma ~ U(3.3, 4.2) A, bond ~ U(1.2, 2.4) A

Guest of the first two: two identical sites, eps 98 K, sigma 3.75 A, bond
1.54 A, equal masses; cutoff 12.8 A, T = 298 K.
"""
from __future__ import annotations

import itertools

import numpy as np

from .cell import cell_matrix

SMALL_SEED = 20261101
FULL_SEED = 20261102
RAGGED_SEED = 20261103

_SHIFTS27 = np.array(list(itertools.product((-1, 0, 1), repeat=3)), dtype=float)


def make_case(lengths, angles, frac, sigma, epsilon, guest_epsilon=98.0, guest_sigma=3.75,
              bond=1.54, site_mass=15.0, cutoff=12.8, temperature=298.0) -> dict:
    half = bond / 2.0
    return {"cell": cell_matrix(lengths, angles), "frame_frac": np.mod(np.asarray(frac, float), 1),
            "frame_sigma": np.asarray(sigma, float), "frame_epsilon": np.asarray(epsilon, float),
            "guest_xyz": np.array([[-half, 0.0, 0.0], [half, 0.0, 0.0]]),
            "guest_sigma": np.array([guest_sigma, guest_sigma], float),
            "guest_epsilon": np.array([guest_epsilon, guest_epsilon], float),
            "guest_mass": np.array([site_mass, site_mass], float),
            "total_mass": 2.0 * site_mass, "temperature": float(temperature),
            "cutoff": float(cutoff), "cell_lengths": [float(v) for v in lengths],
            "cell_angles_deg": [float(v) for v in angles], "bond_length": float(bond)}


def random_frame(rng, cell, n, min_dist=1.2, max_tries=200000):
    """Uniform fractional coordinates, each accepted only if its minimum image
    distance to every accepted atom is >= min_dist."""
    frac = np.empty((n, 3))
    count = tries = 0
    while count < n:
        tries += 1
        if tries > max_tries:
            raise RuntimeError("could not place atoms at the requested minimum distance")
        f = rng.random(3)
        if count and min_dist > 0:
            d = f[None] - frac[:count]
            d -= np.round(d)
            cart = (d[:, None, :] + _SHIFTS27[None]) @ cell
            if np.sqrt((cart ** 2).sum(-1)).min() < min_dist:
                continue
        frac[count] = f
        count += 1
    return frac


def small_field_case(seed=SMALL_SEED) -> dict:
    rng = np.random.default_rng(seed)
    lengths, angles = (25.0, 25.0, 12.0), (90.0, 90.0, 120.0)
    cell = cell_matrix(lengths, angles)
    frac = random_frame(rng, cell, 300)
    types = rng.integers(0, 3, size=300)
    eps = np.array([28.0, 35.0, 8.0])[types]
    sig = np.array([3.40, 3.26, 2.85])[types]
    return make_case(lengths, angles, frac, sig, eps)


def full_field_case(seed=FULL_SEED, n_atoms=1200) -> dict:
    """The full-size framework (``n_atoms`` < 1200 gives a porous variant of
    the same cell and LJ types)."""
    rng = np.random.default_rng(seed)
    lengths, angles = (37.0, 37.0, 14.0), (90.0, 90.0, 120.0)
    cell = cell_matrix(lengths, angles)
    frac = random_frame(rng, cell, n_atoms)
    type_eps = rng.uniform(8.0, 60.0, size=4)
    type_sig = rng.uniform(2.5, 3.8, size=4)
    types = rng.integers(0, 4, size=n_atoms)
    return make_case(lengths, angles, frac, type_sig[types], type_eps[types])


def ragged_cases(n_frameworks=100, guests_per_framework=2, seed=RAGGED_SEED) -> list[dict]:
    rng = np.random.default_rng(seed)
    cases = []
    for _ in range(n_frameworks):
        while True:
            lengths = rng.uniform(10.0, 60.0, size=3)
            angles = rng.uniform(60.0, 120.0, size=3)
            try:
                cell = cell_matrix(lengths, angles)
                break
            except ValueError:
                continue
        n = int(round(np.exp(rng.uniform(np.log(200), np.log(4000)))))
        frac = rng.random((n, 3))
        type_eps = rng.uniform(8.0, 60.0, size=4)
        type_sig = rng.uniform(2.5, 3.8, size=4)
        types = rng.integers(0, 4, size=n)
        for _ in range(guests_per_framework):
            cases.append(make_case(lengths, angles, frac, type_sig[types], type_eps[types],
                                   guest_epsilon=rng.uniform(60.0, 250.0),
                                   guest_sigma=rng.uniform(3.3, 4.2),
                                   bond=rng.uniform(1.2, 2.4)))
    return cases


def toy_case() -> dict:
    """The one-atom skew cell of the reference's examples/toy_triclinic/input.dat
    (12, 13, 14 A; 90, 115.152, 90 deg; rc 6 A; guest eps 108 K, sigma 3.76 A,
    sites at -/+0.77 A, masses 15)."""
    case = make_case((12.0, 13.0, 14.0), (90.0, 115.152, 90.0), [[0.1, 0.2, 0.3]], [3.4], [50.0],
                     guest_epsilon=108.0, guest_sigma=3.76, bond=1.54, cutoff=6.0)
    return case
