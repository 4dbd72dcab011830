"""A guest whose two sites coincide (bond length 0, both offsets 0): every
orientation puts both sites at the center, so each pose energy is twice the
single-site energy there, the same for all directors, and dE/du vanishes.
The kernels' symmetric forms take their o = 0 branches (unscaled director,
zero product-form terms).  Checked against the FP64 oracle and against those
invariants."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def case():
    from string_vext_amd import synthetic
    c = dict(synthetic.small_field_case())
    c["guest_xyz"] = np.zeros((2, 3))
    c["bond_length"] = 0.0
    return c


def test_offsets_are_zero(oracle, case):
    assert np.array_equal(oracle.body_offsets(case), [0.0, 0.0])


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_field_is_twice_the_site_field_for_every_director(oracle, case, precision):
    from string_vext_amd.field import Plan
    size, k = 7, 64
    ref = oracle.vext_fields(case, size, k)
    _, axes = oracle.regular_pose_grid(size, k)
    plan = Plan.from_case(case)
    out = plan.field(plan.grid(size, axes), precision=precision, orientation=True,
                     marginal=True, grid=True, site=True, temperature=case["temperature"])
    e = out["energy"]
    m = oracle.error_metrics(ref["energy"], e)
    if precision == "fp64":
        assert m["all_max_relative_error_floor1K"] < 1e-11
    else:
        assert m["accessible_max_abs_error_K"] < 1e-3
        assert m["all_max_relative_error_floor100K"] < 1e-5
    # the same energy for every director of a center
    spread = np.abs(e - e[:, :1]).max(axis=1) / np.maximum(np.abs(e[:, 0]), 1.0)
    assert spread.max() < (1e-12 if precision == "fp64" else 1e-6)
    # twice the single-site field (the site channel is FP32)
    site = out["site"].astype(np.float64)
    ok = np.abs(e[:, 0]) < 1e6
    np.testing.assert_allclose(e[ok, 0], 2.0 * site[ok], rtol=2e-6, atol=1e-3)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_pose_gradient_has_no_director_part(oracle, case, precision):
    import torch
    from string_vext_amd.autograd import pose_energy
    from string_vext_amd.field import Plan
    rng = np.random.default_rng(11)
    n = 700
    centers = rng.random((n, 3)) @ np.asarray(case["cell"])
    axes = rng.normal(size=(n, 3))
    axes /= np.linalg.norm(axes, axis=1, keepdims=True)
    plan = Plan.from_case(case)
    c = torch.tensor(centers, device="cuda", requires_grad=True)
    u = torch.tensor(axes, device="cuda", requires_grad=True)
    e = pose_energy(plan, c, u, precision)
    e.sum().backward()
    e_ref, gc_ref, gu_ref = oracle.pose_grad(case, centers, axes)
    assert np.array_equal(gu_ref, np.zeros_like(gu_ref))
    assert torch.equal(u.grad, torch.zeros_like(u.grad))
    scale = np.maximum(np.abs(e_ref), 1.0)
    tol = 1e-11 if precision == "fp64" else 1e-5
    acc = e_ref < 2000
    assert (np.abs(e.detach().cpu().numpy() - e_ref) / scale)[acc].max() < tol
    g = c.grad.cpu().numpy()
    gs = np.maximum(np.abs(gc_ref).max(axis=1, keepdims=True), 1.0)
    assert (np.abs(g - gc_ref) / gs)[acc].max() < (1e-9 if precision == "fp64" else 1e-4)


def test_string_iterations_match_cpu_restatement(oracle, case):
    """Strings of the coincident-site guest: no angle force (dV/d angle = 0),
    and iterations equal the CPU restatement."""
    from oracle import string_oracle as SO
    from string_vext_amd.field import Plan
    from string_vext_amd.strings import StringSystem, straight_strings
    plan = Plan.from_case(case)
    x0 = np.concatenate([straight_strings(plan, d, 1, 12, seed=70 + d) for d in (1, 2, 3)])
    dirs = np.array([1, 2, 3])
    setup = oracle.ReuseSetup(case)
    system = StringSystem(plan, x0, dirs, precision="fp64")
    system.run(3, graph=True)
    ref = x0.copy()
    for it in range(3):
        ref = np.stack([SO.iterate(case, ref[s], int(dirs[s]), it, setup=setup)[0]
                        for s in range(len(ref))])
    path, energy, its, conv = system.state()
    assert (its == 3).all()
    np.testing.assert_allclose(path[..., :3], ref[..., :3], rtol=0, atol=1e-10)
    np.testing.assert_allclose(path[..., 3:], ref[..., 3:], rtol=0, atol=1e-9)
    _, g = SO.energy_grad(case, ref[0], setup)
    assert np.abs(g[:, 3:]).max() == 0.0
