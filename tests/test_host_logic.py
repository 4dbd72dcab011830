"""CPU tests of the package's host-side logic against the oracle restatement."""
import re
from pathlib import Path

import numpy as np
import pytest

from tests import _golden

ROOT = Path(__file__).resolve().parents[1]


def test_cell_matrix_equals_oracle(oracle):
    from string_vext_amd.cell import cell_matrix
    rng = np.random.default_rng(1)
    for _ in range(25):
        lengths, angles = rng.uniform(5, 60, 3), rng.uniform(65, 115, 3)
        try:
            ref = oracle.cell_matrix(lengths, angles)
        except ValueError:
            continue
        assert np.array_equal(cell_matrix(lengths, angles), ref)
    with pytest.raises(ValueError):
        cell_matrix((1, 1, 1), (90, 90, 180))
    with pytest.raises(ValueError):
        cell_matrix((1, 1, 1), (10, 10, 170))


def test_pose_grid_equals_oracle(oracle):
    from string_vext_amd.orientation import regular_pose_grid
    for size, k in ((2, 4), (5, 64), (7, 100)):
        a, b = regular_pose_grid(size, k), oracle.regular_pose_grid(size, k)
        np.testing.assert_array_equal(a[0], b[0])
        np.testing.assert_allclose(a[1], b[1], rtol=0, atol=1e-15)
    with pytest.raises(ValueError):
        regular_pose_grid(1, 64)


@pytest.mark.parametrize("name", _golden.NAMES)
def test_input_round_trip_and_parse_equals_oracle(oracle, tmp_path, name):
    from string_vext_amd.input import parse_string_input, write_string_input
    case, _, _, _ = _golden.load(name)
    path = tmp_path / "input.dat"
    write_string_input(case, path)
    ours, ref = parse_string_input(path), oracle.parse_string_input(path)
    for key in ("cell", "frame_frac", "frame_sigma", "frame_epsilon", "guest_xyz", "guest_sigma",
                "guest_epsilon", "guest_mass"):
        np.testing.assert_array_equal(ours[key], ref[key])
        np.testing.assert_allclose(ours[key], case[key], rtol=1e-12, atol=1e-12)
    for key in ("total_mass", "temperature", "cutoff"):
        assert ours[key] == ref[key] == case[key]


def test_parser_rejects_charged_guest(tmp_path):
    from string_vext_amd.input import parse_string_input, write_string_input
    case, _, _, _ = _golden.load("toy")
    path = tmp_path / "input.dat"
    write_string_input(case, path)
    text = path.read_text().splitlines()
    i = next(n for n, line in enumerate(text) if line.startswith("x y z")) + 1
    parts = text[i].split()
    parts[5] = "0.5"
    text[i] = " ".join(parts)
    path.write_text("\n".join(text) + "\n")
    with pytest.raises(ValueError):
        parse_string_input(path)


@pytest.mark.parametrize("name", _golden.NAMES)
def test_reuse_coefficients_equal_oracle(oracle, name):
    from string_vext_amd.neighbor import ReusePotential
    case, _, _, _ = _golden.load(name)
    ours, ref = ReusePotential(case), oracle.ReuseSetup(case)
    assert ours.radius == ref.radius and ours.cutoff == ref.cutoff
    np.testing.assert_array_equal(ours.offsets, ref.offsets)
    np.testing.assert_array_equal(ours.potential.images, ref.images)
    for key in ("coeff12", "coeff6", "shift", "floor2"):
        np.testing.assert_array_equal(getattr(ours, key), getattr(ref, key))


def test_validate_case_rejections():
    from string_vext_amd.physics import validate_case
    case, _, _, _ = _golden.load("toy")
    validate_case(case)
    for key, value in (("frame_sigma", np.array([-1.0])), ("guest_mass", np.array([1.0, 1.0])),
                       ("cutoff", 0.0), ("guest_xyz", np.array([[0, 1.0, 0], [1, 0, 0]]))):
        bad = dict(case)
        bad[key] = value
        with pytest.raises(ValueError):
            validate_case(bad)


def test_reuse_requires_identical_sites():
    from string_vext_amd.neighbor import ReusePotential
    case, _, _, _ = _golden.load("toy")
    bad = dict(case)
    bad["guest_sigma"] = np.array([3.7, 3.8])
    with pytest.raises(ValueError):
        ReusePotential(bad)


def test_synthetic_cases_are_seeded_and_respect_min_distance():
    from string_vext_amd import synthetic
    a, b = synthetic.small_field_case(), synthetic.small_field_case()
    np.testing.assert_array_equal(a["frame_frac"], b["frame_frac"])
    frac, cell = a["frame_frac"], a["cell"]
    d = frac[:, None] - frac[None]
    d -= np.round(d)
    import itertools
    shifts = np.array(list(itertools.product((-1, 0, 1), repeat=3)), dtype=float)
    cart = (d[:, :, None, :] + shifts) @ cell
    r = np.sqrt((cart ** 2).sum(-1)).min(-1)
    r[np.arange(len(r)), np.arange(len(r))] = np.inf
    assert r.min() >= 1.2
    full = synthetic.full_field_case()
    assert len(full["frame_frac"]) == 1200 and len(np.unique(full["frame_sigma"])) == 4


def test_cli_refuses_existing_output(tmp_path):
    from string_vext_amd.vext import main
    out = tmp_path / "vext.npz"
    out.write_bytes(b"")
    with pytest.raises(FileExistsError):
        main(["--input", str(tmp_path / "missing.dat"), "--output", str(out)])


def test_candidate_csr_equals_ball_queries(oracle):
    """The drop-in route's CSR (sparse distance matrix) holds, for every center,
    the same images as the oracle's cKDTree ball queries (native.py:47-55)."""
    from string_vext_amd.native import candidate_csr
    case, centers, _, _ = _golden.load("c2h4")
    setup = oracle.ReuseSetup(case)
    cart = np.mod(centers, 1) @ setup.cell
    starts, indices = candidate_csr(setup.tree, cart, setup.radius)
    ref_starts, ref_indices = oracle.csr(setup.tree, cart, setup.radius)
    assert starts.dtype == np.int64 and indices.dtype == np.int32
    np.testing.assert_array_equal(starts, ref_starts)
    for c in range(len(cart)):
        assert set(indices[starts[c]:starts[c + 1]]) == set(ref_indices[ref_starts[c]:ref_starts[c + 1]])


def test_pose_energy_route_follows_autograd_recording(monkeypatch):
    """pose_energy's default takes the fused route (gradient in the forward's
    site pass) exactly when autograd records the call; an explicit flag wins."""
    import torch
    from string_vext_amd import autograd
    seen = []
    monkeypatch.setattr(autograd.PoseEnergy, "apply",
                        lambda c, u, plan, precision, fused: seen.append(fused))
    c = torch.zeros(2, 3, dtype=torch.float64, requires_grad=True)
    u = torch.ones(2, 3, dtype=torch.float64)
    autograd.pose_energy(None, c, u)
    autograd.pose_energy(None, c.detach(), u.requires_grad_(True))
    with torch.no_grad():
        autograd.pose_energy(None, c, u)
    autograd.pose_energy(None, c.detach(), u.detach())
    autograd.pose_energy(None, c, u, fused=False)
    autograd.pose_energy(None, c.detach(), u.detach(), fused=True)
    assert seen == [True, True, False, False, False, True]
