"""The allocation-free drop-in entry point (evaluate_reused_neighbors_noalloc)
gives the results and return codes of evaluate_reused_neighbors, bit for bit,
on two calls in a row (the second smaller, so stale scratch contents would
show), and rejects a call whose device staging exceeds its scratch."""
import ctypes

import numpy as np
import pytest

from tests import _golden

pytestmark = pytest.mark.gpu

_P = ctypes.c_void_p


def _call(fn, setup, cart, axes, starts, indices, n_images=None):
    out = np.full((len(cart), len(axes)), np.nan)
    arrays = [np.ascontiguousarray(cart, dtype=np.float64),
              np.ascontiguousarray(axes, dtype=np.float64), np.ascontiguousarray(setup.images),
              np.ascontiguousarray(setup.offsets), np.ascontiguousarray(setup.coeff12),
              np.ascontiguousarray(setup.coeff6), np.ascontiguousarray(setup.shift),
              np.ascontiguousarray(setup.floor2), np.ascontiguousarray(starts, dtype=np.int64),
              np.ascontiguousarray(indices, dtype=np.int32), out]
    rc = fn(len(cart), len(axes), len(setup.images) if n_images is None else n_images,
            setup.cutoff ** 2, *[a.ctypes.data_as(_P) for a in arrays])
    return rc, out


@pytest.fixture(scope="module")
def lib():
    from string_vext_amd import _lib
    return _lib.load()


def test_noalloc_equals_dropin_on_two_calls(oracle, lib):
    from string_vext_amd import synthetic
    case = synthetic.small_field_case()
    setup = oracle.ReuseSetup(case)
    centers, axes = oracle.regular_pose_grid(6, 64)
    cart = np.mod(centers, 1) @ setup.cell
    for n in (len(cart), 37):                      # the second call is smaller
        starts, indices = oracle.csr(setup.tree, cart[:n], setup.radius)
        rc_a, a = _call(lib.evaluate_reused_neighbors, setup, cart[:n], axes, starts, indices)
        rc_b, b = _call(lib.evaluate_reused_neighbors_noalloc, setup, cart[:n], axes, starts,
                        indices)
        assert rc_a == rc_b == 0
        assert np.array_equal(a, b)


def test_noalloc_matches_reference_kernel_bits(oracle, lib):
    case, centers, axes, ref = _golden.load("c2h4")
    setup = oracle.ReuseSetup(case)
    axes = axes / np.linalg.norm(axes, axis=1, keepdims=True)
    cart = np.mod(centers, 1) @ setup.cell
    starts, indices = oracle.csr(setup.tree, cart, setup.radius)
    rc, out = _call(lib.evaluate_reused_neighbors_noalloc, setup, cart, axes, starts, indices)
    assert rc == 0 and np.array_equal(out, ref)


def test_noalloc_return_codes_match_dropin(oracle, lib):
    """Codes 2 and 3 stop at the first bad center, whose predecessors are computed."""
    from string_vext_amd import synthetic
    case = synthetic.small_field_case()
    setup = oracle.ReuseSetup(case)
    centers, axes = oracle.regular_pose_grid(3, 8)
    cart = np.mod(centers, 1) @ setup.cell
    starts, indices = oracle.csr(setup.tree, cart, setup.radius)
    bad_starts = starts.copy()
    bad_starts[5] = bad_starts[6] + 1               # center 5 decreases
    bad_index = indices.copy()
    bad_index[starts[3]] = len(setup.images)        # center 3 holds an out-of-range index
    for s, i, code, stop in ((bad_starts, indices, 2, 5), (starts, bad_index, 3, 3)):
        rc_a, a = _call(lib.evaluate_reused_neighbors, setup, cart, axes, s, i)
        rc_b, b = _call(lib.evaluate_reused_neighbors_noalloc, setup, cart, axes, s, i)
        assert rc_a == rc_b == code
        assert np.array_equal(a[:stop], b[:stop])


def test_noalloc_rejects_staging_beyond_its_scratch(lib):
    """1.5 M images need about 84 MB of staging, more than the 64 MiB scratch:
    rejected with code 1 before any device work."""
    n_images = 1_500_000

    class Big:
        images = np.zeros((n_images, 3))
        offsets = np.array([-0.77, 0.77])
        coeff12 = coeff6 = shift = floor2 = np.ones(n_images)
        cutoff = 12.8

    cart = np.zeros((1, 3))
    axes = np.array([[1.0, 0.0, 0.0]])
    starts = np.array([0, 0], dtype=np.int64)
    indices = np.zeros(0, dtype=np.int32)
    rc, _ = _call(lib.evaluate_reused_neighbors_noalloc, Big, cart, axes, starts, indices)
    assert rc == 1
    assert b"scratch" in lib.sv_last_error()
