"""libsv_vext.so loads on the CPU host and exports every entry point of
include/sv_vext.h (no compute call is made without a GPU)."""
import ctypes
import re
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
HEADER = ROOT / "include" / "sv_vext.h"


def declared_functions():
    text = re.sub(r"/\*.*?\*/", "", HEADER.read_text(), flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(\w+)\s*\(", text, flags=re.M)))


def test_header_declares_the_dropin_and_batched_api():
    names = declared_functions()
    assert "evaluate_reused_neighbors" in names
    for name in ("sv_plan_create", "sv_poses_grid", "sv_field_run", "sv_batch_run"):
        assert name in names


def test_library_exports_every_declared_symbol():
    from string_vext_amd import _lib
    handle = ctypes.CDLL(str(_lib.LIB_PATH))
    missing = [n for n in declared_functions() if not hasattr(handle, n)]
    assert not missing, missing


def test_binding_covers_header():
    from string_vext_amd import _lib
    assert sorted(n for n, _, _ in _lib.SIGNATURES) == declared_functions()


def test_library_reports_no_device_on_cpu_host():
    from string_vext_amd import _lib
    lib = _lib.load(require_device=False)
    assert lib.sv_abi_version() == 5
    if lib.sv_device_count() == 0:
        with pytest.raises(RuntimeError):
            _lib.load(require_device=True)


def test_dropin_validates_sizes_before_touching_the_device():
    """Return code 1 of kernel.cpp:11-12 needs no GPU."""
    from string_vext_amd import _lib
    lib = _lib.load(require_device=False)
    null = None
    assert lib.evaluate_reused_neighbors(-1, 4, 1, 1.0, *([null] * 11)) == 1
    assert lib.evaluate_reused_neighbors(1, 0, 1, 1.0, *([null] * 11)) == 1
    assert lib.evaluate_reused_neighbors(1, 4, 0, 1.0, *([null] * 11)) == 1
    assert lib.evaluate_reused_neighbors(1, 4, 1, 0.0, *([null] * 11)) == 1
    assert lib.evaluate_reused_neighbors(0, 4, 1, 1.0, *([null] * 11)) == 0


def test_argument_validation_codes_on_cpu_host():
    """Invalid arguments are rejected with code 1 before any device work; a valid
    plan on a host without a GPU fails with code 5 (HIP runtime)."""
    import numpy as np
    from string_vext_amd import _lib
    lib = _lib.load(require_device=False)
    P = ctypes.c_void_p

    def ptr(a):
        return a.ctypes.data_as(P)

    frac = np.zeros((1, 3))
    sig, eps, off = np.array([3.0]), np.array([10.0]), np.array([-0.77, 0.77])
    h = P()

    def plan(cell, n=1, sigma=sig, epsilon=eps, cutoff=12.0):
        cell = np.ascontiguousarray(cell, dtype=np.float64).ravel()
        return lib.sv_plan_create(ptr(cell), n, ptr(frac), ptr(sigma), ptr(epsilon), 3.0, 100.0,
                                  ptr(off), cutoff, 50.0, ctypes.byref(h))

    assert plan(np.diag([10.0, 10.0, -10.0])) == 1            # det <= 0
    assert plan(np.eye(3) * 10, n=0) == 1
    assert plan(np.eye(3) * 10, sigma=-sig) == 1
    assert plan(np.eye(3) * 10, epsilon=np.array([np.nan])) == 1
    assert plan(np.eye(3) * 10, cutoff=0.0) == 1
    assert plan(np.eye(3) * 10) == 5                           # no device here
    assert lib.sv_pose_grid_eval(None, 1, None, None, 0, None, None, None, None) == 1
    assert lib.sv_pose_grid_check(None) == 1
    assert lib.sv_strings_create(0, None, 5, None, None, 1e-4, 0.01, 0, ctypes.byref(h)) == 1
    assert lib.sv_batch_run(None, 0, 1) == 1
    assert b"null" in lib.sv_last_error() or lib.sv_last_error()
