"""Guests with unequal site masses: the center of mass is off the bond middle,
so the site offsets are not symmetric (o1 != -o0) and every kernel takes its
general two-offset form.  Checked against the FP64 oracle."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def case():
    from string_vext_amd import synthetic
    c = dict(synthetic.small_field_case())
    c["guest_mass"] = np.array([13.0, 17.0])
    c["total_mass"] = 30.0
    return c


def test_offsets_are_asymmetric(oracle, case):
    o = oracle.body_offsets(case)
    assert abs(o[0] + o[1]) > 0.1


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_field_all_channels(oracle, case, precision):
    from string_vext_amd.field import Plan
    size, k = 7, 64
    ref = oracle.vext_fields(case, size, k)
    _, axes = oracle.regular_pose_grid(size, k)
    plan = Plan.from_case(case)
    out = plan.field(plan.grid(size, axes), precision=precision, orientation=True,
                     marginal=True, grid=True, site=True, temperature=case["temperature"])
    m = oracle.error_metrics(ref["energy"], out["energy"])
    if precision == "fp64":
        # rounding level (the FP64 form differs from the oracle only in the order
        # of its additions and in tile-relative coordinates)
        assert m["all_max_relative_error_floor1K"] < 1e-11
    else:
        assert m["accessible_max_abs_error_K"] < 1e-3
        # absolute errors here are ~1e-4 K; a pose with |E| of a few K then reads
        # 2e-5 against the 1 K floor
        assert m["all_max_relative_error_floor1K"] < 1e-4
    np.testing.assert_allclose(out["marginal"].reshape(ref["marginal_K"].shape),
                               ref["marginal_K"], rtol=1e-6, atol=1e-3)
    np.testing.assert_allclose(out["site"].reshape(ref["site_K"][0].shape), ref["site_K"][0],
                               rtol=1e-6, atol=1e-3)


def test_pose_grid_energy_and_gradient(oracle, case):
    import torch
    from string_vext_amd.field import Plan, PoseGrid
    rng = np.random.default_rng(31)
    n = 400
    centers = rng.random((n, 3)) @ np.asarray(case["cell"])
    axes = rng.normal(size=(n, 3))
    axes /= np.linalg.norm(axes, axis=1, keepdims=True)
    e_ref, gc_ref, gu_ref = oracle.pose_grad(case, centers, axes)
    grid = PoseGrid(Plan.from_case(case), n_expected=n)
    c = torch.tensor(centers, dtype=torch.float64, device="cuda")
    u = torch.tensor(axes, dtype=torch.float64, device="cuda")
    for precision, tol_e, tol_g in (("fp64", 1e-12, 1e-10), ("mixed", 1e-5, 1e-4)):
        e = torch.empty(n, dtype=torch.float64, device="cuda")
        gc = torch.empty((n, 3), dtype=torch.float64, device="cuda")
        gu = torch.empty((n, 3), dtype=torch.float64, device="cuda")
        grid.eval(n, c.data_ptr(), u.data_ptr(), precision, energy=e.data_ptr(),
                  grad_center=gc.data_ptr(), grad_axis=gu.data_ptr())
        e = e.cpu().numpy()
        assert (np.abs(e - e_ref) / np.maximum(np.abs(e_ref), 1.0)).max() < tol_e
        for g, ref in ((gc.cpu().numpy(), gc_ref), (gu.cpu().numpy(), gu_ref)):
            scale = np.maximum(np.abs(ref).max(axis=1, keepdims=True), 1.0)
            assert (np.abs(g - ref) / scale).max() < tol_g


@pytest.mark.parametrize("persistent", [0, 5])
def test_batch_with_symmetric_and_asymmetric_guests(oracle, case, persistent):
    from string_vext_amd import synthetic
    from string_vext_amd.field import Batch, Plan
    cases = [case, synthetic.small_field_case()]
    jobs, refs = [], []
    for c, size in zip(cases, (5, 6)):
        plan = Plan.from_case(c)
        centers, axes = oracle.regular_pose_grid(size, 64)
        jobs.append((plan, plan.grid(size, axes)))
        refs.append(oracle.reuse_energy(c, centers, axes))
    batch = Batch(jobs, precision="mixed", outputs=("energy",))
    if persistent:
        batch.set_persistent(persistent)
    batch.run(sync=True)
    for j in range(len(jobs)):
        m = oracle.error_metrics(refs[j], batch.fetch("energy", j))
        assert m["all_max_relative_error_floor100K"] < 1e-5
        if m["accessible_poses"]:
            assert m["accessible_max_abs_error_K"] < 1e-3


def test_string_iterations_match_cpu_restatement(oracle, case):
    from oracle import string_oracle as SO
    from string_vext_amd.field import Plan
    from string_vext_amd.strings import StringSystem, straight_strings
    plan = Plan.from_case(case)
    x0 = straight_strings(plan, 2, 1, 10, seed=5)
    setup = oracle.ReuseSetup(case)
    system = StringSystem(plan, x0, np.array([2]), precision="fp64")
    system.run(3, graph=True)
    ref = x0[0].copy()
    for it in range(3):
        ref = SO.iterate(case, ref, 2, it, setup=setup)[0]
    path = system.state()[0][0]
    np.testing.assert_allclose(path[:, :3], ref[:, :3], rtol=0, atol=1e-10)
    np.testing.assert_allclose(path[:, 3:], ref[:, 3:], rtol=0, atol=1e-9)
