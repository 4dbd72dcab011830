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
    assert lib.sv_abi_version() == 3
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
