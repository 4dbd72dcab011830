"""Pins the CPU restatement of one String iteration (oracle/string_oracle.py)
by the invariants the reference's kernels guarantee."""
import numpy as np

from oracle import string_oracle as SO


def test_smoothing_closed_form_and_wrap_branches():
    rng = np.random.default_rng(2)
    ang = rng.uniform(1.0, 2.0, size=(7, 3))          # no wrapping anywhere
    out = SO.smooth(ang)
    np.testing.assert_array_equal(out[0], ang[0])
    np.testing.assert_array_equal(out[-1], ang[-1])
    expect = (1 - 1e-4) * ang[1:-1] + 0.5e-4 * (ang[:-2] + ang[2:])
    np.testing.assert_allclose(out[1:-1], expect, rtol=1e-15)
    # neighbours across the 0/2pi seam: left at 6.2, centre 0.1, right 0.2
    a = np.array([[6.2, 0, 0], [0.1, 0, 0], [0.2, 0, 0]])
    got = SO.smooth(a)[1, 0]
    assert abs(got - ((1 - 1e-4) * 0.1 + 0.5e-4 * (6.2 + 0.2 - 2 * np.pi))) < 1e-15


def test_step_invariants_on_a_synthetic_gradient(monkeypatch):
    """Endpoint fix, periodic last image and arc-length reparametrization."""
    cell = np.diag([10.0, 12.0, 9.0])
    case = {"cell": cell}
    n = 9
    rng = np.random.default_rng(4)
    s0 = np.zeros((n, 6))
    s0[:, 0] = np.linspace(0, 1, n)
    s0[:, 1:3] = 0.3 + 0.01 * rng.random((n, 2))
    s0[:, 3:] = rng.uniform(0.5, 1.0, (n, 3))
    grad = rng.normal(size=(n, 6)) * 50.0
    grad[:, 3] = 0.0
    monkeypatch.setattr(SO, "energy_grad",
                        lambda case_, x, setup=None: (np.zeros(len(x)), grad.copy()))
    out, conv = SO.iterate(case, s0, 1, it=1)
    assert not conv
    assert out[0, 0] == s0[0, 0] and out[-1, 0] == s0[-1, 0]
    np.testing.assert_array_equal(out[-1, 1:3], out[0, 1:3])
    cart = out[:, :3] @ cell
    seg = np.sqrt(((cart[1:] - cart[:-1]) ** 2).sum(1))
    # equal arc-length points along a nearly straight polyline: equal chords
    np.testing.assert_allclose(seg, seg.mean(), rtol=2e-3)
    # the normalized translational step has length move_frac in fractional space
    mv = 1e-4
    g = SO._floor(mv * grad)
    nt = SO._floor(np.sqrt((g[:, :3] ** 2).sum(1)))
    assert (nt >= mv).all()


def test_convergence_thresholds():
    n, mf, ma = 96, 1e-4, np.radians(1.0)
    assert np.isclose(0.30 * mf * n * np.sqrt(3), 0.004988, rtol=1e-3)
    assert np.isclose(0.30 * ma * n * np.sqrt(3), 0.870568, rtol=1e-3)
