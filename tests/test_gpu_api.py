"""The reference's DIRECT/site API on the GPU (src/string_vext/physics.py:59-106,
133-164) against the oracle, and the reference's own differential test of the
hot path (tests/unit/test_public_api.py:52-62) repeated through this package:
calculate_vext(size=3, orientations=4) on the toy skew cell against
direct_pose_energy at rtol 1e-6, atol 1e-3, with site_K of shape (2, 3, 3, 3)."""
from pathlib import Path

import numpy as np
import pytest

from tests import _golden

pytestmark = pytest.mark.gpu

TOY = Path(__file__).resolve().parent / "golden" / "toy_triclinic_input.dat"


@pytest.mark.parametrize("name", _golden.NAMES)
def test_direct_pose_energy_matches_oracle(oracle, name):
    from string_vext_amd.physics import direct_pose_energy
    case, centers, axes, ref = _golden.load(name)
    e = direct_pose_energy(case, centers, axes)
    direct = oracle.direct_pose_energy(case, centers, axes)
    assert e.shape == (len(centers), len(axes))
    np.testing.assert_allclose(e, direct, rtol=1e-10, atol=1e-7)
    # and the reference kernel's own outputs on the same poses (kernel.cpp golden)
    np.testing.assert_allclose(e, ref, rtol=1e-10, atol=1e-4)


@pytest.mark.parametrize("name", _golden.NAMES)
def test_site_potential_matches_oracle(oracle, name):
    from string_vext_amd.physics import TriclinicSitePotential
    case, centers, _, _ = _golden.load(name)
    rng = np.random.default_rng(7)
    cell = np.asarray(case["cell"])
    # points inside and outside the cell (the potential wraps them)
    points = np.concatenate([centers @ cell, (rng.random((50, 3)) * 3 - 1) @ cell])
    pot = TriclinicSitePotential(case)
    for sigma, eps in ((case["guest_sigma"][0], case["guest_epsilon"][0]), (3.3, 150.0)):
        e = pot.evaluate(points, sigma, eps)
        ref = oracle.site_energy(case, points, sigma, eps)
        np.testing.assert_allclose(e, ref, rtol=1e-10, atol=1e-7)


def test_uniform_fields_match_oracle(oracle):
    from string_vext_amd import synthetic
    from string_vext_amd.physics import uniform_fields
    case = synthetic.small_field_case()
    size, k = 4, 64
    out = uniform_fields(case, size, k)
    ref = oracle.vext_fields(case, size, k)
    assert out["orientation_K"].shape == (k, size, size, size)
    assert out["site_K"].shape == (2, size, size, size)
    np.testing.assert_allclose(out["orientation_K"], ref["orientation_K"], rtol=1e-6, atol=1e-3)
    np.testing.assert_allclose(out["marginal_K"], ref["marginal_K"], rtol=1e-6, atol=1e-3)
    np.testing.assert_allclose(out["site_K"], ref["site_K"], rtol=1e-6, atol=1e-3)
    np.testing.assert_array_equal(out["axes"], ref["axes"])
    with pytest.raises(NotImplementedError):
        uniform_fields(case, size, k, energy_mode="interpolate")


def test_public_api_native_vext_matches_direct_on_toy_skew_cell():
    """test_public_api.py:52-62 through this package, on the reference's own
    examples/toy_triclinic/input.dat."""
    from string_vext_amd.input import parse_string_input
    from string_vext_amd.orientation import regular_pose_grid
    from string_vext_amd.physics import direct_pose_energy
    from string_vext_amd.vext import calculate_vext
    case = parse_string_input(TOY)
    fields = calculate_vext(case, size=3, orientations=4)
    centers, axes = regular_pose_grid(3, 4)
    direct = direct_pose_energy(case, centers, axes)
    np.testing.assert_allclose(fields["orientation_K"].reshape(4, -1).T, direct, rtol=1e-6,
                               atol=1e-3)
    assert fields["site_K"].shape == (2, 3, 3, 3)
