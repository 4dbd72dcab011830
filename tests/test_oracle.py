"""Pins the FP64 CPU oracle before it is trusted as the checker.

* the reference kernel's own outputs (golden fixtures from kernel.cpp compiled
  from source) and, where built, oracle/_ref/kernel.so itself -- bit for bit;
* a brute-force sum over an oversized block of lattice images (no neighbor
  list, no reach formula);
* the DIRECT formula at the reference test's tolerances
  (tests/unit/test_public_api.py:60-61);
* closed forms: cell KAT (test_public_api.py:43-50), pose-grid directions,
  mixing rules, marginal;
* the String pair restatement against the reference's header compiled through
  its host shim and against the brute force of test_native_string_triclinic.py;
* gradients against central differences.
Planted mistakes (in the oracle, on the CPU) show that each check can fail.
"""
import itertools

import numpy as np
import pytest

from tests import _golden


@pytest.mark.parametrize("name", _golden.NAMES)
def test_oracle_reproduces_reference_kernel_outputs_bitwise(oracle, name):
    case, centers, axes, ref = _golden.load(name)
    out = oracle.reuse_energy(case, centers, axes)
    assert np.array_equal(out, ref)


def test_oracle_equals_reference_kernel_on_random_poses(oracle):
    fn = oracle.reference_kernel()
    if fn is None:
        pytest.skip("oracle/_ref/kernel.so not built (reference tree absent)")
    from string_vext_amd import synthetic
    case = synthetic.small_field_case()
    rng = np.random.default_rng(11)
    centers = rng.random((40, 3))
    axes = rng.normal(size=(16, 3))
    setup = oracle.ReuseSetup(case)
    a = oracle.reuse_energy(case, centers, axes, setup=setup)
    b = oracle.reuse_energy(case, centers, axes, setup=setup, kernel="reference")
    assert np.array_equal(a, b)


def test_planted_shift_error_is_caught(oracle):
    """A dropped cutoff shift in the oracle must break parity with the reference bits."""
    case, centers, axes, ref = _golden.load("c2h4")
    setup = oracle.ReuseSetup(case)
    setup.shift = np.zeros_like(setup.shift)
    out = oracle.reuse_energy(case, centers, axes, setup=setup)
    assert np.abs(out - ref).max() > 1.0


def test_planted_missing_image_is_caught(oracle):
    """Dropping the nearest lattice shell of images must break the brute-force check."""
    case, centers, axes, _ = _golden.load("c2h4")
    axes = axes / np.linalg.norm(axes, axis=1, keepdims=True)
    full = oracle.brute_pose_energy(case, centers, axes, reach=4)
    short = oracle.brute_pose_energy(case, centers, axes, reach=1)
    assert np.abs(full - short).max() > 1e-3


@pytest.mark.parametrize("name", _golden.NAMES)
def test_reuse_equals_brute_force_images(oracle, name):
    case, centers, axes, ref = _golden.load(name)
    cell = np.asarray(case["cell"])
    spacing = 1.0 / np.linalg.norm(np.linalg.inv(cell), axis=0)
    reach = int(np.ceil((case["cutoff"] + 2.0) / spacing.min())) + 1
    brute = oracle.brute_pose_energy(case, centers, axes, reach=reach)
    np.testing.assert_allclose(ref, brute, rtol=1e-9, atol=1e-7)


def test_direct_matches_reuse_like_public_api(oracle):
    """tests/unit/test_public_api.py:52-62 and test_exact_neighbor_reuse_native.py:15-33."""
    case, _, _, _ = _golden.load("toy")
    centers, axes = oracle.regular_pose_grid(3, 4)
    reuse = oracle.reuse_energy(case, centers, axes)
    direct = oracle.direct_pose_energy(case, centers, axes)
    np.testing.assert_allclose(reuse, direct, rtol=1e-6, atol=1e-3)
    for name in ("c2h4", "c2h6"):
        case, centers, axes, ref = _golden.load(name)
        direct = oracle.direct_pose_energy(case, centers, axes)
        np.testing.assert_allclose(ref, direct, rtol=1e-10, atol=1e-4)


def test_cell_matrix_known_answer(oracle):
    """test_public_api.py:43-50: beta = 115.152 deg gives cell[2,0] < 0, |c| = 14."""
    cell = oracle.cell_matrix((12.0, 13.0, 14.0), (90.0, 115.152, 90.0))
    assert cell[2, 0] < 0
    assert abs(np.linalg.norm(cell[2]) - 14.0) < 1e-12
    rng = np.random.default_rng(3)
    for _ in range(20):
        lengths = rng.uniform(5, 50, 3)
        angles = rng.uniform(70, 110, 3)
        m = oracle.cell_matrix(lengths, angles)
        np.testing.assert_allclose(np.linalg.norm(m, axis=1), lengths, rtol=1e-12)
        a, b, c = m
        cos = lambda u, v: u @ v / np.linalg.norm(u) / np.linalg.norm(v)  # noqa: E731
        got = np.degrees(np.arccos([cos(b, c), cos(a, c), cos(a, b)]))
        np.testing.assert_allclose(got, angles, rtol=1e-10)
        assert m[0, 1] == m[0, 2] == m[1, 2] == 0.0


def test_pose_grid_closed_form(oracle):
    centers, axes = oracle.regular_pose_grid(4, 64)
    assert centers.shape == (64, 3) and axes.shape == (64, 3)
    np.testing.assert_allclose(centers[1], [0.125, 0.125, 0.375])   # ic fastest
    np.testing.assert_allclose(centers[16], [0.375, 0.125, 0.125])  # ia slowest
    np.testing.assert_allclose(np.linalg.norm(axes, axis=1), 1.0, rtol=1e-15)
    np.testing.assert_allclose(axes[:, 2], 1 - (2 * np.arange(64) + 1) / 64)
    golden = np.pi * (3 - np.sqrt(5))
    phi = np.unwrap(np.arctan2(axes[:, 1], axes[:, 0]))
    np.testing.assert_allclose(np.diff(phi) % (2 * np.pi), golden % (2 * np.pi), atol=1e-12)


def test_mixing_rule_closed_form(oracle):
    case, _, _, _ = _golden.load("c2h4")
    setup = oracle.ReuseSetup(case)
    n = len(case["frame_frac"])
    s = (case["frame_sigma"][0] + case["guest_sigma"][0]) / 2
    e = np.sqrt(case["frame_epsilon"][0] * case["guest_epsilon"][0])
    assert setup.coeff12[0] == pytest.approx(4 * e * s ** 12, rel=1e-14)
    assert setup.coeff6[0] == pytest.approx(4 * e * s ** 6, rel=1e-14)
    assert setup.floor2[0] == pytest.approx((0.1 * s) ** 2, rel=1e-14)
    # the shifted pair term vanishes at the cutoff (shift taken at rc, not R)
    rc = case["cutoff"]
    at_rc = setup.coeff12[0] / rc ** 12 - setup.coeff6[0] / rc ** 6 + setup.shift[0]
    assert abs(at_rc) < 1e-12
    assert setup.images.shape[0] % n == 0


def test_marginal_closed_form(oracle):
    e = np.full((3, 64), -123.25)
    np.testing.assert_allclose(oracle.marginal_energy(e, 298.0), -123.25, rtol=1e-14)
    two = np.array([[0.0, 100.0]])
    expect = -298.0 * np.log((1 + np.exp(-100 / 298.0)) / 2)
    np.testing.assert_allclose(oracle.marginal_energy(two, 298.0), expect, rtol=1e-14)


def _brute_pair(cell, site, atom, sigma=3.5, epsilon=98.0, cutoff=5.5):
    """test_native_string_triclinic.py:41-49: oversized image block, shifts -5..5."""
    shift = np.array(list(itertools.product(range(-5, 6), repeat=3)))
    guest = ((site @ np.linalg.inv(cell)) % 1) @ cell
    images = ((atom % 1)[None] + shift) @ cell
    r = np.linalg.norm(guest - images, axis=1)
    s = (sigma + 3.4) / 2
    e = np.sqrt(epsilon * 50.0)
    ratio = (s / np.maximum(r, 0.1 * s)) ** 6
    cut = (s / cutoff) ** 6
    return np.where(np.maximum(r, 0.1 * s) < cutoff, 4 * e * (ratio ** 2 - ratio - cut ** 2 + cut),
                    0.0).sum()


def _pair(oracle, cell, site, atom, cutoff=5.5):
    return oracle.string_pair(site, atom, cell, 98.0, 3.5, 50.0, 3.4, cutoff)


def test_string_pair_matches_reference_header(oracle):
    fn = oracle.reference_string_pair()
    if fn is None:
        pytest.skip("oracle/_ref/string_pair.so not built (reference tree absent)")
    import ctypes
    rng = np.random.default_rng(20260929)
    for h in (np.diag([9.0, 10.0, 11.0]), np.array([[10.0, 0, 0], [4.0, 8.0, 0], [-3.0, 2.0, 9.0]])):
        for _ in range(32):
            f, atom = rng.random((2, 3))
            site = f @ h
            p = np.r_[site, atom, h.T.reshape(-1), 98.0, 3.5, 50.0, 3.4, 5.5].astype(np.float64)
            ref = fn(p.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
            assert _pair(oracle, h, site, atom) == ref


def test_string_pair_brute_force_cases(oracle):
    """The reference's invariants (test_native_string_triclinic.py:59-90)."""
    h = np.array([[10.0, 0.0, 0.0], [8.0, 6.0, 0.0], [0.0, 0.0, 10.0]])
    site, atom = np.array([0.0, 0.0, 5.0]), np.array([0.0, 0.9, 0.5])
    e = _pair(oracle, h, site, atom, cutoff=4)
    assert e == pytest.approx(_brute_pair(h, site, atom, cutoff=4), abs=1e-5) and e > 1e6
    rng = np.random.default_rng(20260929)
    for h in (np.diag([9.0, 10.0, 11.0]), np.array([[10.0, 0, 0], [4.0, 8.0, 0], [-3.0, 2.0, 9.0]])):
        for _ in range(32):
            f, atom = rng.random((2, 3))
            np.testing.assert_allclose(_pair(oracle, h, f @ h, atom), _brute_pair(h, f @ h, atom),
                                       rtol=2e-10, atol=1e-6)
    h = np.diag([4.0, 4.0, 4.0])
    atom, site = np.array([0.2, 0.3, 0.4]), np.array([2.0, 2.0, 2.0])
    np.testing.assert_allclose(_pair(oracle, h, site, atom, 6), _brute_pair(h, site, atom, cutoff=6),
                               rtol=2e-10)
    h = np.array([[10.0, 0, 0], [4.0, 8.0, 0], [-3.0, 2.0, 9.0]])
    f, atom = np.array([0.25, 0.35, 0.6]), np.array([0.8, 0.1, 0.2])
    v = _pair(oracle, h, f @ h, atom)
    for shift in (np.array([3, -2, 1]), np.array([-5, 1, -4])):
        np.testing.assert_allclose(v, _pair(oracle, h, (f + shift) @ h, atom), rtol=1e-10, atol=1e-7)
        np.testing.assert_allclose(v, _pair(oracle, h, f @ h, atom + shift), rtol=1e-10, atol=1e-7)
    h = np.diag([20.0, 20.0, 20.0])
    atom = np.array([0.5, 0.5, 0.5])
    e = _pair(oracle, h, atom @ h, atom)
    assert np.isfinite(e) and e > 1e12
    assert _pair(oracle, h, atom @ h + [5.5, 0, 0], atom) == pytest.approx(0.0, abs=1e-10)


def test_pose_gradient_matches_central_differences(oracle):
    from string_vext_amd import synthetic
    case = synthetic.small_field_case()
    setup = oracle.ReuseSetup(case)
    rng = np.random.default_rng(5)
    centers = rng.random((12, 3)) @ setup.cell
    axes = rng.normal(size=(12, 3))
    axes /= np.linalg.norm(axes, axis=1, keepdims=True)
    e, gc, gu = oracle.pose_grad(case, centers, axes, setup=setup)
    ref_e = np.array([oracle.reuse_energy(case, np.linalg.solve(setup.cell.T, c)[None], u[None],
                                          setup=setup)[0, 0] for c, u in zip(centers, axes)])
    np.testing.assert_allclose(e, ref_e, rtol=1e-12, atol=1e-9)
    h = 1e-5
    for k in range(3):
        dc = np.zeros(3)
        dc[k] = h
        ep = oracle.pose_grad(case, centers + dc, axes, setup=setup)[0]
        em = oracle.pose_grad(case, centers - dc, axes, setup=setup)[0]
        fd = (ep - em) / (2 * h)
        np.testing.assert_allclose(gc[:, k], fd, rtol=1e-5, atol=1e-4 * (1 + np.abs(fd).max()))
        ep = oracle.pose_grad(case, centers, axes + dc, setup=setup)[0]
        em = oracle.pose_grad(case, centers, axes - dc, setup=setup)[0]
        fd = (ep - em) / (2 * h)
        np.testing.assert_allclose(gu[:, k], fd, rtol=1e-5, atol=1e-4 * (1 + np.abs(fd).max()))


def test_error_metrics_floors(oracle):
    """The 1 K floor is the reference probe's (exact_neighbor_reuse_probe.py:73-83);
    the 100 K floor turns a 1e-5 bound into 1e-3 K absolute below 100 K."""
    ref = np.array([-5.0, 0.5, 250.0, 3.0e4])
    got = ref + np.array([5e-4, 2e-5, 1e-3, 0.2])
    m = oracle.error_metrics(ref, got)
    assert m["accessible_poses"] == 3
    assert m["accessible_max_abs_error_K"] == pytest.approx(1e-3)
    assert m["all_max_relative_error_floor1K"] == pytest.approx(1e-4)     # 5e-4 / 5
    assert m["all_max_relative_error_floor100K"] == pytest.approx(2e-5 / 3)  # 0.2 / 3e4
    assert m["all_max_abs_error_K"] == pytest.approx(0.2)
