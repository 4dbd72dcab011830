"""GPU parity of independent-pose energies and their hand-written gradient
(sv_pose_list_*, string_vext_amd.autograd) against the FP64 oracle, central
finite differences, and the String solver's forward-difference forces."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def case():
    from string_vext_amd import synthetic
    return synthetic.small_field_case()


@pytest.fixture(scope="module")
def poses(case):
    rng = np.random.default_rng(20261104)
    n = 700
    frac = rng.random((n, 3))
    frac[:8] = [[.001, .999, .5], [.999, .001, .5], [.5, .001, .999], [.001, .5, .999],
                [.999, .5, .001], [.5, .999, .001], [.001, .001, .001], [.999, .999, .999]]
    axes = rng.normal(size=(n, 3))
    axes /= np.linalg.norm(axes, axis=1, keepdims=True)
    return frac @ np.asarray(case["cell"]), axes


def _plan(case):
    from string_vext_amd.field import Plan
    return Plan.from_case(case)


def test_pose_energy_and_gradient_fp64_match_oracle(oracle, case, poses):
    from string_vext_amd.field import PoseList
    centers, axes = poses
    e_ref, gc_ref, gu_ref = oracle.pose_grad(case, centers, axes)
    pl = PoseList(_plan(case), centers, axes)
    e, gc, gu = pl.energy_host("fp64", grad=True)
    scale = np.maximum(np.abs(e_ref), 1.0)
    assert (np.abs(e - e_ref) / scale).max() < 1e-12
    gscale = np.maximum(np.abs(gc_ref).max(axis=1, keepdims=True), 1.0)
    assert (np.abs(gc - gc_ref) / gscale).max() < 1e-10
    gscale = np.maximum(np.abs(gu_ref).max(axis=1, keepdims=True), 1.0)
    assert (np.abs(gu - gu_ref) / gscale).max() < 1e-10


def test_pose_energy_and_gradient_mixed_within_tolerance(oracle, case, poses):
    from string_vext_amd.field import PoseList
    centers, axes = poses
    e_ref, gc_ref, gu_ref = oracle.pose_grad(case, centers, axes)
    e, gc, gu = PoseList(_plan(case), centers, axes).energy_host("mixed", grad=True)
    m = oracle.error_metrics(e_ref, e)
    assert m["accessible_max_abs_error_K"] < 1e-3
    assert m["all_max_relative_error_floor1K"] < 1e-5
    # gradient: K/A; relative to the pose's gradient scale with a 1 K/A floor
    for g, ref in ((gc, gc_ref), (gu, gu_ref)):
        scale = np.maximum(np.abs(ref).max(axis=1, keepdims=True), 1.0)
        assert (np.abs(g - ref) / scale).max() < 1e-4


def test_gradient_matches_central_differences(case, poses):
    """Central differences with h = 1e-4 A of the GPU FP64 energies."""
    from string_vext_amd.field import PoseList
    centers, axes = poses
    centers, axes = centers[:120], axes[:120]
    plan = _plan(case)
    e, gc, gu = PoseList(plan, centers, axes).energy_host("fp64", grad=True)
    h = 1e-4
    for k in range(3):
        step = np.zeros(3)
        step[k] = h
        ep = PoseList(plan, centers + step, axes).energy_host("fp64")
        em = PoseList(plan, centers - step, axes).energy_host("fp64")
        fd = (ep - em) / (2 * h)
        ok = np.abs(e) < 1e4    # central differences of overlap poses are not informative
        np.testing.assert_allclose(gc[ok, k], fd[ok], rtol=1e-5, atol=1e-4)
        ep = PoseList(plan, centers, axes + step).energy_host("fp64")
        em = PoseList(plan, centers, axes - step).energy_host("fp64")
        fd = (ep - em) / (2 * h)
        np.testing.assert_allclose(gu[ok, k], fd[ok], rtol=1e-5, atol=1e-4)


def test_string_solver_energy_and_forward_difference_forces(oracle, case):
    """The String solver's pose energy (triclinic_pair.h summed over sites and
    atoms, body-x guest rotated by Rz(g)Ry(b)Rx(a), device_fxn.h:34-63) and its
    unnormalized forward-difference forces g_k = V(x + d e_k) - V(x) with the
    input's deltas (global.h:1281-1324; d_frac = 1e-4, d_angle = 1 deg) against
    the analytic gradient by the chain rule."""
    from string_vext_amd.field import PoseList
    cell = np.asarray(case["cell"])
    off = oracle.body_offsets(case)
    gs, ge = case["guest_sigma"][0], case["guest_epsilon"][0]
    rng = np.random.default_rng(9)

    def director(beta, gamma):
        return np.array([np.cos(gamma) * np.cos(beta), np.sin(gamma) * np.cos(beta), -np.sin(beta)])

    def v_string(x):
        c = x[:3] @ cell
        u = director(x[4], x[5])
        total = 0.0
        for o in off:
            site = c + o * u
            for f, s, e in zip(case["frame_frac"], case["frame_sigma"], case["frame_epsilon"]):
                total += oracle.string_pair(site, f, cell, ge, gs, e, s, case["cutoff"])
        return total

    plan = _plan(case)
    # accessible poses (E < 2000 K) found by a GPU scan of random poses
    xs_all = np.c_[rng.random((4000, 3)), rng.uniform(0, 2 * np.pi, (4000, 3))]
    dirs = np.array([director(b, g) for b, g in xs_all[:, 4:6]])
    scan = PoseList(plan, xs_all[:, :3] @ cell, dirs).energy_host("fp64")
    picks = xs_all[(scan < 2000.0) & (scan > -1e4)][:6]
    assert len(picks) >= 3
    tried = 0
    for x in picks:
        c, u = x[:3] @ cell, director(x[4], x[5])
        e, gc, gu = PoseList(plan, c[None], u[None]).energy_host("fp64", grad=True)
        v = v_string(x)
        tried += 1
        assert e[0] == pytest.approx(v, rel=1e-10, abs=1e-7)
        deltas = [1e-4, 1e-4, 1e-4, np.radians(1.0), np.radians(1.0), np.radians(1.0)]
        for k in (0, 1, 2, 4, 5):
            xs = x.copy()
            xs[k] += deltas[k]
            g_fd = v_string(xs) - v
            # analytic derivative at the midpoint: forward difference is second-order there
            xm = x.copy()
            xm[k] += deltas[k] / 2
            cm, um = xm[:3] @ cell, director(xm[4], xm[5])
            _, gcm, gum = PoseList(plan, cm[None], um[None]).energy_host("fp64", grad=True)
            if k < 3:
                deriv = (cell @ gcm[0])[k]
            elif k == 4:
                deriv = gum[0] @ np.array([-np.cos(xm[5]) * np.sin(xm[4]),
                                           -np.sin(xm[5]) * np.sin(xm[4]), -np.cos(xm[4])])
            else:
                deriv = gum[0] @ np.array([-np.sin(xm[5]) * np.cos(xm[4]),
                                           np.cos(xm[5]) * np.cos(xm[4]), 0.0])
            # the midpoint-rule residual of a forward difference is f''' d^3 / 24:
            # ~1e-5 relative for d = 1e-4 fractional (2.5e-3 A), ~1e-3 for 1 deg
            tol = 1e-4 if k < 3 else 5e-3
            assert g_fd == pytest.approx(deriv * deltas[k], rel=tol, abs=1e-6)
    assert tried >= 3


def test_autograd_function(oracle, case, poses):
    import torch
    from string_vext_amd.autograd import pose_energy
    centers, axes = poses
    e_ref, gc_ref, gu_ref = oracle.pose_grad(case, centers, axes)
    plan = _plan(case)
    c = torch.tensor(centers, dtype=torch.float64, device="cuda", requires_grad=True)
    u = torch.tensor(axes, dtype=torch.float64, device="cuda", requires_grad=True)
    w = torch.linspace(0.5, 1.5, len(centers), dtype=torch.float64, device="cuda")
    e = pose_energy(plan, c, u, "fp64")
    (e * w).sum().backward()
    wn = w.cpu().numpy()[:, None]
    np.testing.assert_allclose(e.detach().cpu().numpy(), e_ref, rtol=1e-12, atol=1e-9)
    np.testing.assert_allclose(c.grad.cpu().numpy(), wn * gc_ref, rtol=1e-9,
                               atol=1e-9 * np.abs(gc_ref).max())
    np.testing.assert_allclose(u.grad.cpu().numpy(), wn * gu_ref, rtol=1e-9,
                               atol=1e-9 * np.abs(gu_ref).max())
    # normalized directors as free parameters: autograd carries the normalization
    v = torch.tensor(axes * 2.5, dtype=torch.float64, device="cuda", requires_grad=True)
    e2 = pose_energy(plan, c.detach(), v / v.norm(dim=1, keepdim=True), "mixed")
    e2.sum().backward()
    assert torch.isfinite(v.grad).all()
    radial = (v.grad * v.detach()).sum(dim=1).abs().max().item()
    assert radial < 1e-6 * max(1.0, v.grad.abs().max().item())


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_autograd_fused_gradient_equals_two_pass(oracle, case, poses, precision):
    """pose_energy(..., fused=True): energy and gradient from one site pass in the
    forward, scaled by the upstream gradient in the backward.  Same site sums as
    the energy kernel and the recomputing backward kernel, so the results are
    those of the two-pass route; FP64 also against the oracle's gradient."""
    import torch
    from string_vext_amd.autograd import check_poses, pose_energy
    centers, axes = poses
    plan = _plan(case)
    w = torch.linspace(0.5, 1.5, len(centers), dtype=torch.float64, device="cuda")
    out = {}
    for fused in (False, True):
        c = torch.tensor(centers, dtype=torch.float64, device="cuda", requires_grad=True)
        u = torch.tensor(axes, dtype=torch.float64, device="cuda", requires_grad=True)
        e = pose_energy(plan, c, u, precision, fused=fused)
        (e * w).sum().backward()
        check_poses(plan)
        out[fused] = [t.detach().cpu().numpy() for t in (e, c.grad, u.grad)]
    for a, b in zip(out[False], out[True]):
        print(f"fused vs two-pass max abs diff {np.abs(b - a).max():.3e}", flush=True)
        np.testing.assert_allclose(b, a, rtol=1e-12, atol=1e-12 * np.abs(a).max())
    if precision == "fp64":
        e_ref, gc_ref, gu_ref = oracle.pose_grad(case, centers, axes)
        wn = w.cpu().numpy()[:, None]
        np.testing.assert_allclose(out[True][0], e_ref, rtol=1e-12, atol=1e-9)
        np.testing.assert_allclose(out[True][1], wn * gc_ref, rtol=1e-9,
                                   atol=1e-9 * np.abs(gc_ref).max())
        np.testing.assert_allclose(out[True][2], wn * gu_ref, rtol=1e-9,
                                   atol=1e-9 * np.abs(gu_ref).max())
    # no input needs a gradient: the fused flag runs the energy kernel alone
    with torch.no_grad():
        e0 = pose_energy(plan, torch.tensor(centers, device="cuda"),
                         torch.tensor(axes, device="cuda"), precision, fused=True)
    np.testing.assert_array_equal(e0.cpu().numpy(), out[False][0])


def test_pose_grid_device_binning_matches_oracle_and_host_binning(oracle, case, poses):
    """sv_pose_grid_eval: poses binned on the GPU from device arrays, against the
    FP64 oracle (energy and gradient) and against the host-binned pose list."""
    import torch
    from string_vext_amd.field import PoseGrid, PoseList
    centers, axes = poses
    n = len(centers)
    e_ref, gc_ref, gu_ref = oracle.pose_grad(case, centers, axes)
    plan = _plan(case)
    grid = PoseGrid(plan, n_expected=n)
    c = torch.tensor(centers, dtype=torch.float64, device="cuda")
    u = torch.tensor(axes, dtype=torch.float64, device="cuda")
    for precision in ("fp64", "mixed"):
        e = torch.empty(n, dtype=torch.float64, device="cuda")
        gc = torch.empty((n, 3), dtype=torch.float64, device="cuda")
        gu = torch.empty((n, 3), dtype=torch.float64, device="cuda")
        grid.eval(n, c.data_ptr(), u.data_ptr(), precision, energy=e.data_ptr())
        grid.eval(n, c.data_ptr(), u.data_ptr(), precision, grad_center=gc.data_ptr(),
                  grad_axis=gu.data_ptr())
        e, gc, gu = e.cpu().numpy(), gc.cpu().numpy(), gu.cpu().numpy()
        if precision == "fp64":
            assert (np.abs(e - e_ref) / np.maximum(np.abs(e_ref), 1.0)).max() < 1e-12
            for g, ref in ((gc, gc_ref), (gu, gu_ref)):
                scale = np.maximum(np.abs(ref).max(axis=1, keepdims=True), 1.0)
                assert (np.abs(g - ref) / scale).max() < 1e-10
            e_host = PoseList(plan, centers, axes).energy_host("fp64")
            assert (np.abs(e - e_host) / np.maximum(np.abs(e_host), 1.0)).max() < 1e-13
        else:
            m = oracle.error_metrics(e_ref, e)
            assert m["accessible_max_abs_error_K"] < 1e-3
            assert m["all_max_relative_error_floor1K"] < 1e-5
            for g, ref in ((gc, gc_ref), (gu, gu_ref)):
                scale = np.maximum(np.abs(ref).max(axis=1, keepdims=True), 1.0)
                assert (np.abs(g - ref) / scale).max() < 1e-4


def test_pose_grid_grows_shrinks_and_rejects_nonfinite_poses(oracle, case, poses):
    import torch
    from string_vext_amd.field import PoseGrid
    centers, axes = poses
    plan = _plan(case)
    grid = PoseGrid(plan, n_expected=64)
    for n in (len(centers), 5, 300):          # grow past the initial capacity, then reuse
        c = torch.tensor(centers[:n], dtype=torch.float64, device="cuda")
        u = torch.tensor(axes[:n], dtype=torch.float64, device="cuda")
        e = torch.empty(n, dtype=torch.float64, device="cuda")
        grid.eval(n, c.data_ptr(), u.data_ptr(), "fp64", energy=e.data_ptr())
        e_ref = oracle.pose_grad(case, centers[:n], axes[:n])[0]
        assert (np.abs(e.cpu().numpy() - e_ref) / np.maximum(np.abs(e_ref), 1.0)).max() < 1e-12
    assert grid.info()["capacity"] >= len(centers)
    # a nonfinite center is flagged by the binning kernel and reported as code 1
    bad = torch.tensor(centers[:8], dtype=torch.float64, device="cuda")
    bad[3, 1] = float("nan")
    u = torch.tensor(axes[:8], dtype=torch.float64, device="cuda")
    e = torch.empty(8, dtype=torch.float64, device="cuda")
    with pytest.raises(RuntimeError, match="code 1"):
        grid.eval(8, bad.data_ptr(), u.data_ptr(), "mixed", energy=e.data_ptr())
    # and the grid stays usable
    c = torch.tensor(centers[:8], dtype=torch.float64, device="cuda")
    grid.eval(8, c.data_ptr(), u.data_ptr(), "fp64", energy=e.data_ptr())
    assert np.isfinite(e.cpu().numpy()).all()


def test_pose_grid_async_on_torch_stream_and_flag_check(oracle, case, poses):
    """sv_pose_grid_eval_async enqueues on the caller's stream with no host
    wait; its results equal the synchronous call, and invalid input is
    reported later by sv_pose_grid_check."""
    import torch
    from string_vext_amd.autograd import check_poses, pose_energy
    from string_vext_amd.field import PoseGrid
    centers, axes = poses
    n = len(centers)
    plan = _plan(case)
    grid = PoseGrid(plan, n_expected=n)
    c = torch.tensor(centers, dtype=torch.float64, device="cuda")
    u = torch.tensor(axes, dtype=torch.float64, device="cuda")
    e_sync = torch.empty(n, dtype=torch.float64, device="cuda")
    grid.eval(n, c.data_ptr(), u.data_ptr(), "mixed", energy=e_sync.data_ptr())
    side = torch.cuda.Stream()
    with torch.cuda.stream(side):
        e_async = torch.empty(n, dtype=torch.float64, device="cuda")
        grid.eval_async(side.cuda_stream, n, c.data_ptr(), u.data_ptr(), "mixed",
                        energy=e_async.data_ptr())
        doubled = e_async * 2.0            # ordered after the evaluation on the same stream
    side.synchronize()
    grid.check()
    assert torch.equal(e_async, e_sync)
    assert torch.equal(doubled, 2.0 * e_sync)
    bad = c.clone()
    bad[5, 0] = float("inf")
    e_bad = torch.empty(n, dtype=torch.float64, device="cuda")
    grid.eval_async(torch.cuda.current_stream().cuda_stream, n, bad.data_ptr(), u.data_ptr(),
                    "fp64", energy=e_bad.data_ptr())
    with pytest.raises(RuntimeError, match="code 1"):
        grid.check()
    grid.check()                            # flags were cleared
    # autograd path: no host wait; invalid poses show as NaN and in check_poses
    e = pose_energy(plan, bad, u, "fp64")
    assert torch.isnan(e[5]) or not torch.isfinite(e[5])
    with pytest.raises(RuntimeError, match="code 1"):
        check_poses(plan)


def test_pose_grid_reeval_reuses_binning(case, poses):
    """sv_pose_grid_reeval_async evaluates the last binned poses again without
    reading the caller's arrays: bit-identical to a call that bins them anew.
    autograd's backward takes that route only while the grid still holds its
    forward's poses; after another forward on the grid it bins again."""
    import torch
    from string_vext_amd import _lib
    from string_vext_amd.autograd import _grid_for, pose_energy
    from string_vext_amd.field import PoseGrid
    centers, axes = poses
    n = len(centers)
    plan = _plan(case)
    grid = PoseGrid(plan, n_expected=n)
    stream = torch.cuda.current_stream().cuda_stream
    with pytest.raises(_lib.VextError, match="code 1"):
        grid.reeval_async(stream, "fp64", energy=0)
    c = torch.tensor(centers, dtype=torch.float64, device="cuda")
    u = torch.tensor(axes, dtype=torch.float64, device="cuda")
    for precision in ("fp64", "mixed"):
        e = torch.empty(n, dtype=torch.float64, device="cuda")
        ref = [torch.empty((n, 3), dtype=torch.float64, device="cuda") for _ in range(2)]
        got = [torch.empty((n, 3), dtype=torch.float64, device="cuda") for _ in range(2)]
        g = torch.linspace(-1.0, 2.0, n, dtype=torch.float64, device="cuda")
        grid.eval_async(stream, n, c.data_ptr(), u.data_ptr(), precision, grad_out=g.data_ptr(),
                        grad_center=ref[0].data_ptr(), grad_axis=ref[1].data_ptr())
        grid.eval_async(stream, n, c.data_ptr(), u.data_ptr(), precision, energy=e.data_ptr())
        c_saved = c.clone()
        c.fill_(float("nan"))                # the binned copies are used, not c
        grid.reeval_async(stream, precision, grad_out=g.data_ptr(),
                          grad_center=got[0].data_ptr(), grad_axis=got[1].data_ptr())
        grid.check()
        c.copy_(c_saved)
        assert torch.equal(got[0], ref[0]) and torch.equal(got[1], ref[1])
    # autograd, two-pass route: forward A, forward B on the same grid, then backward of A
    cg = torch.tensor(centers, dtype=torch.float64, device="cuda", requires_grad=True)
    ug = torch.tensor(axes, dtype=torch.float64, device="cuda", requires_grad=True)
    ea = pose_energy(plan, cg, ug, "fp64", fused=False)
    eb = pose_energy(plan, cg.detach().flip(0), ug.detach().flip(0), "fp64", fused=False)
    assert _grid_for(plan, n).binned_token is not None
    ea.sum().backward()
    gc_a, gu_a = cg.grad.clone(), ug.grad.clone()
    cg.grad = None
    ug.grad = None
    pose_energy(plan, cg, ug, "fp64", fused=False).sum().backward()      # reuses the binning
    assert torch.equal(cg.grad, gc_a) and torch.equal(ug.grad, gu_a)
    assert torch.equal(eb.flip(0), ea.detach())


def test_pose_grid_census_matches_oracle(oracle, case, poses):
    """The exact in-cutoff (site, image) pair count of independent poses."""
    import torch
    from string_vext_amd.field import PoseGrid
    centers, axes = poses
    centers, axes = centers[:200], axes[:200]
    plan = _plan(case)
    grid = PoseGrid(plan, len(centers))
    c = torch.tensor(centers, device="cuda")
    u = torch.tensor(axes, device="cuda")
    e = torch.empty(len(c), dtype=torch.float64, device="cuda")
    grid.eval(len(c), c.data_ptr(), u.data_ptr(), "fp64", energy=e.data_ptr())
    frac = centers @ np.linalg.inv(np.asarray(case["cell"]))
    setup = oracle.ReuseSetup(case)
    ref = sum(oracle.count_pairs(case, f[None], a[None], setup=setup) for f, a in zip(frac, axes))
    assert grid.census() == ref


def test_non_unit_directors_are_used_as_given(oracle, case, poses):
    """Sites at c + o u for |u| != 1, identically in both precisions: mixed and
    FP64 energies and gradients against the oracle's (candidates taken within
    rc + max|o| max|u|)."""
    from string_vext_amd.field import PoseList
    centers, axes = poses
    rng = np.random.default_rng(31)
    scaled = axes * rng.uniform(0.5, 1.7, size=(len(axes), 1))
    setup = oracle.ReuseSetup(case)
    setup.radius = case["cutoff"] + np.abs(setup.offsets).max() * 1.7
    e_ref, gc_ref, gu_ref = oracle.pose_grad(case, centers, scaled, setup=setup)
    plan = _plan(case)
    for precision, e_tol, g_tol in (("fp64", 1e-12, 1e-10), ("mixed", None, 1e-4)):
        e, gc, gu = PoseList(plan, centers, scaled).energy_host(precision, grad=True)
        if e_tol:
            assert (np.abs(e - e_ref) / np.maximum(np.abs(e_ref), 1.0)).max() < e_tol
        else:
            m = oracle.error_metrics(e_ref, e)
            assert m["all_max_relative_error_floor100K"] < 1e-5
            if m["accessible_poses"]:
                assert m["accessible_max_abs_error_K"] < 1e-3
        for g, ref in ((gc, gc_ref), (gu, gu_ref)):
            scale = np.maximum(np.abs(ref).max(axis=1, keepdims=True), 1.0)
            assert (np.abs(g - ref) / scale).max() < g_tol


@pytest.mark.parametrize("fused", [True, False])
def test_second_derivative_is_refused(case, poses, fused):
    """The kernels give first derivatives only: differentiating the gradient
    again raises instead of returning zeros."""
    import torch
    from string_vext_amd.autograd import pose_energy
    centers, axes = poses
    plan = _plan(case)
    c = torch.tensor(centers[:64], dtype=torch.float64, device="cuda", requires_grad=True)
    u = torch.tensor(axes[:64], dtype=torch.float64, device="cuda")
    e = pose_energy(plan, c, u, "fp64", fused=fused)
    # a loss whose gradient in E depends on c: the second derivative would need
    # the Hessian of E, which no kernel provides
    (g,) = torch.autograd.grad((e * e).sum(), c, create_graph=True)
    assert torch.isfinite(g).all()
    with pytest.raises(RuntimeError, match="once_differentiable"):
        g.sum().backward()


@pytest.mark.parametrize("fused", [True, False])
def test_host_tensors_get_host_energies_and_gradients(case, poses, fused):
    """Host (CPU) inputs are evaluated on the GPU; their energies and gradients
    come back on the host, equal to those of the same inputs on the GPU."""
    import torch
    from string_vext_amd.autograd import pose_energy
    centers, axes = poses
    plan = _plan(case)
    out = {}
    for device in ("cpu", "cuda"):
        c = torch.tensor(centers[:200], dtype=torch.float64, device=device, requires_grad=True)
        u = torch.tensor(axes[:200], dtype=torch.float64, device=device, requires_grad=True)
        e = pose_energy(plan, c, u, "mixed", fused=fused)
        assert e.device.type == device
        e.sum().backward()
        assert c.grad.device.type == device and u.grad.device.type == device
        out[device] = [t.detach().cpu() for t in (e, c.grad, u.grad)]
    for a, b in zip(out["cpu"], out["cuda"]):
        assert torch.equal(a, b)
