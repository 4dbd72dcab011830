"""Edge cases of every entry point, against the FP64 oracle: centers whose
candidate lists are empty, zero centers, a one-atom framework in a cell
smaller than the cutoff (hundreds of images of one atom), single poses,
poses exactly on the cell boundary and coincident poses, and the shortest
string (3 images)."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

_P = ctypes.c_void_p


@pytest.fixture(scope="module")
def lib():
    from string_vext_amd import _lib
    return _lib.load()


@pytest.fixture(scope="module")
def small_case():
    from string_vext_amd import synthetic
    return synthetic.small_field_case()


@pytest.fixture(scope="module")
def one_atom_case():
    """One atom in a 7 A triclinic cell with a 12.8 A cutoff."""
    from string_vext_amd import synthetic
    return synthetic.make_case((7.0, 7.5, 8.0), (80.0, 95.0, 110.0), [[0.3, 0.6, 0.1]], [3.3],
                               [40.0])


def _abi(lib, name, setup, cart, axes, starts, indices):
    out = np.full((max(len(cart), 1), len(axes)), np.nan)
    arrays = [np.ascontiguousarray(cart, dtype=np.float64).reshape(-1, 3),
              np.ascontiguousarray(axes, dtype=np.float64), np.ascontiguousarray(setup.images),
              np.ascontiguousarray(setup.offsets), np.ascontiguousarray(setup.coeff12),
              np.ascontiguousarray(setup.coeff6), np.ascontiguousarray(setup.shift),
              np.ascontiguousarray(setup.floor2), np.ascontiguousarray(starts, dtype=np.int64),
              np.ascontiguousarray(indices, dtype=np.int32).reshape(-1), out]
    rc = getattr(lib, name)(len(cart), len(axes), len(setup.images), setup.cutoff ** 2,
                            *[a.ctypes.data_as(_P) for a in arrays])
    return rc, out[:len(cart)]


@pytest.mark.parametrize("name", ["evaluate_reused_neighbors", "evaluate_reused_neighbors_noalloc"])
def test_dropin_empty_candidate_lists_and_zero_centers(oracle, lib, small_case, name):
    """Rows whose CSR range is empty are exact zeros, as in the reference
    kernel's loop (kernel.cpp:13-44); zero centers return 0 and write nothing."""
    setup = oracle.ReuseSetup(small_case)
    centers, axes = oracle.regular_pose_grid(3, 16)
    cart = np.mod(centers, 1) @ setup.cell
    starts, indices = oracle.csr(setup.tree, cart, setup.radius)
    # empty every third center's candidate range
    keep = np.ones(len(cart), bool)
    keep[::3] = False
    lengths = np.diff(starts) * keep
    pieces = [indices[starts[i]:starts[i + 1]] for i in range(len(cart)) if keep[i]]
    idx2 = np.concatenate(pieces).astype(np.int32)
    starts2 = np.concatenate([[0], np.cumsum(lengths)]).astype(np.int64)
    rc, out = _abi(lib, name, setup, cart, axes, starts2, idx2)
    assert rc == 0
    ref = oracle.run_reuse_kernel(oracle.lib().oracle_reuse, cart, axes, setup, starts2, idx2)
    assert np.array_equal(out[::3], np.zeros_like(out[::3]))
    np.testing.assert_allclose(out, ref, rtol=1e-12, atol=1e-12)
    # every range empty
    rc, out = _abi(lib, name, setup, cart, axes, np.zeros(len(cart) + 1, np.int64),
                   np.zeros(1, np.int32))
    assert rc == 0 and np.array_equal(out, np.zeros_like(out))
    # zero centers
    rc, _ = _abi(lib, name, setup, np.zeros((0, 3)), axes, np.zeros(1, np.int64),
                 np.zeros(1, np.int32))
    assert rc == 0


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_one_atom_framework_in_a_cell_smaller_than_the_cutoff(oracle, one_atom_case, precision):
    from string_vext_amd.field import Plan
    case = one_atom_case
    setup = oracle.ReuseSetup(case)
    assert len(setup.images) > 200              # one atom, many periodic images
    centers, axes = oracle.regular_pose_grid(5, 16)
    ref = oracle.reuse_energy(case, centers, axes, setup=setup)
    plan = Plan.from_case(case)
    e = plan.field(plan.poses(centers, axes), precision=precision)["energy"]
    m = oracle.error_metrics(ref, e)
    if precision == "fp64":
        assert (np.abs(e - ref) / np.maximum(np.abs(ref), 1.0)).max() < 1e-10
    else:
        assert m["all_max_relative_error_floor100K"] < 1e-5
        if m["accessible_poses"]:
            assert m["accessible_max_abs_error_K"] < 1e-3
    assert plan.census(plan.poses(centers, axes)) == oracle.count_pairs(case, centers, axes,
                                                                       setup=setup)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_single_boundary_and_coincident_poses(oracle, small_case, precision):
    """The pose grid with one pose, then with poses on the cell's faces and
    corners (fractional 0 and 1, which wrap to the same place), a pose repeated
    three times, and a pose outside the cell: energies and gradients against
    the oracle, and equal results for the poses that coincide."""
    import torch
    from string_vext_amd.autograd import pose_energy
    from string_vext_amd.field import Plan
    case = small_case
    cell = np.asarray(case["cell"])
    plan = Plan.from_case(case)
    setup = oracle.ReuseSetup(case)
    u0 = np.array([[0.6, -0.48, 0.64]])
    frac = np.array([[0.0, 0.0, 0.0], [1.0, 1.0, 1.0], [1.0, 0.0, 0.5], [0.0, 1.0, 0.5],
                     [0.37, 0.61, 0.25], [0.37, 0.61, 0.25], [0.37, 0.61, 0.25],
                     [1.37, -0.39, 2.25]])
    axes = np.repeat(u0, len(frac), axis=0)
    for sel in (slice(4, 5), slice(None)):
        c = frac[sel] @ cell
        u = axes[sel]
        e_ref, gc_ref, gu_ref = oracle.pose_grad(case, c, u, setup=setup)
        cc = torch.tensor(c, device="cuda", requires_grad=True)
        uu = torch.tensor(u, device="cuda", requires_grad=True)
        e = pose_energy(plan, cc, uu, precision)
        e.sum().backward()
        e = e.detach().cpu().numpy()
        gc, gu = cc.grad.cpu().numpy(), uu.grad.cpu().numpy()
        tol = 1e-10 if precision == "fp64" else 1e-5
        assert (np.abs(e - e_ref) / np.maximum(np.abs(e_ref), 100.0)).max() < tol
        for g, ref in ((gc, gc_ref), (gu, gu_ref)):
            scale = np.maximum(np.abs(ref).max(axis=1, keepdims=True), 1.0)
            assert (np.abs(g - ref) / scale).max() < (1e-9 if precision == "fp64" else 1e-4)
        if len(e) == len(frac):
            # faces and corners that are the same point, and the repeated pose
            assert abs(e[0] - e[1]) <= 1e-9 * max(1.0, abs(e[0]))
            assert abs(e[2] - e[3]) <= 1e-9 * max(1.0, abs(e[2]))
            assert e[4] == e[5] == e[6]
            # a pose one lattice vector (and more) away is the same pose
            assert abs(e[7] - e[4]) <= 1e-9 * max(1.0, abs(e[4]))


def test_shortest_string_matches_cpu_restatement(oracle, small_case):
    from oracle import string_oracle as SO
    from string_vext_amd.field import Plan
    from string_vext_amd.strings import StringSystem, straight_strings
    plan = Plan.from_case(small_case)
    x0 = straight_strings(plan, 2, 1, 3, seed=9, scan=6, orientations=4, probes=6)
    setup = oracle.ReuseSetup(small_case)
    system = StringSystem(plan, x0, np.array([2]), precision="fp64")
    ref = x0[0]
    for it in range(3):
        ref = SO.iterate(small_case, ref, 2, it, setup=setup)[0]
    system.run(3, graph=False)
    path, energy, its, _ = system.state()
    assert its[0] == 3
    np.testing.assert_allclose(path[0, :, :3], ref[:, :3], rtol=0, atol=1e-10)
    np.testing.assert_allclose(path[0, :, 3:], ref[:, 3:], rtol=0, atol=1e-9)
    np.testing.assert_allclose(energy[0], SO.energy_grad(small_case, ref, setup)[0], rtol=1e-10,
                               atol=1e-7)
