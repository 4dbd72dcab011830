"""ctypes binding of libsv_vext.so (the C ABI of include/sv_vext.h).

The library is the product: nothing in this package computes energies on the
CPU.  ``load()`` raises when the library is missing or no GPU is visible, so a
missing build fails loudly instead of falling back.
"""
from __future__ import annotations

import ctypes
import threading
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "libsv_vext.so"

_P = ctypes.c_void_p
_I = ctypes.c_int
_D = ctypes.c_double
_lock = threading.Lock()
_lib = None

# (name, restype, argtypes) of every entry point of include/sv_vext.h
SIGNATURES = [
    ("evaluate_reused_neighbors", _I, [_I, _I, _I, _D] + [_P] * 11),
    ("evaluate_reused_neighbors_noalloc", _I, [_I, _I, _I, _D] + [_P] * 11),
    ("sv_last_error", ctypes.c_char_p, []),
    ("sv_abi_version", _I, []),
    ("sv_device_count", _I, []),
    ("sv_plan_create", _I, [_P, _I, _P, _P, _P, _D, _D, _P, _D, _D, _P]),
    ("sv_plan_destroy", _I, [_P]),
    ("sv_plan_info", _I, [_P, _P, _P]),
    ("sv_poses_create", _I, [_P, _I, _P, _I, _P, _P]),
    ("sv_poses_grid", _I, [_P, _I, _I, _P, _P]),
    ("sv_poses_destroy", _I, [_P]),
    ("sv_poses_create_ex", _I, [_P, _I, _P, _I, _P, _I, _P]),
    ("sv_poses_grid_ex", _I, [_P, _I, _I, _P, _I, _P]),
    ("sv_field_run", _I, [_P, _P, _I, _D, _P, _P, _P, _P, _P]),
    ("sv_field_census", _I, [_P, _P, _P]),
    ("sv_batch_create", _I, [_I, _P, _P, _I, _P, _P, _P, _P, _P, _P, _P]),
    ("sv_batch_run", _I, [_P, _I, _I]),
    ("sv_batch_sync", _I, [_P]),
    ("sv_batch_timings", _I, [_P, _P]),
    ("sv_batch_census", _I, [_P, _P]),
    ("sv_batch_destroy", _I, [_P]),
    ("sv_batch_set_persistent", _I, [_P, _I]),
    ("sv_pose_list_create", _I, [_P, _I, _P, _P, _P]),
    ("sv_pose_list_eval", _I, [_P, _I, _P, _P, _P, _P]),
    ("sv_pose_grid_create", _I, [_P, _I, _I, _P]),
    ("sv_pose_grid_eval", _I, [_P, _I, _P, _P, _I, _P, _P, _P, _P]),
    ("sv_pose_grid_info", _I, [_P, _P, _P]),
    ("sv_pose_grid_destroy", _I, [_P]),
    ("sv_pose_grid_eval_async", _I, [_P, _P, _I, _P, _P, _I, _P, _P, _P, _P]),
    ("sv_pose_grid_check", _I, [_P]),
    ("sv_pose_grid_reeval_async", _I, [_P, _P, _I, _P, _P, _P, _P]),
    ("sv_pose_grid_census", _I, [_P, _P]),
    ("sv_pose_grid_set_timing", _I, [_P, _I]),
    ("sv_pose_grid_set_parts", _I, [_P, _I]),
    ("sv_pose_grid_timings", _I, [_P, _P]),
    ("sv_strings_create", _I, [_I, _P, _I, _P, _P, _D, _D, _I, _P]),
    ("sv_strings_run", _I, [_P, _I, _I]),
    ("sv_strings_state", _I, [_P, _P, _P, _P, _P]),
    ("sv_strings_destroy", _I, [_P]),
    ("sv_device_malloc", _I, [_P, ctypes.c_uint64]),
    ("sv_device_free", _I, [_P]),
    ("sv_memcpy_d2h", _I, [_P, _P, ctypes.c_uint64]),
    ("sv_memcpy_h2d", _I, [_P, _P, ctypes.c_uint64]),
]


class VextError(RuntimeError):
    """A nonzero return code of libsv_vext.so."""

    def __init__(self, code: int, message: str):
        super().__init__(f"libsv_vext failed with code {code}: {message}")
        self.code = code


def _open(path: Path):
    # torch ships its own HIP runtime; when torch is importable it is loaded
    # first so that this library binds to the same runtime instance and device
    # pointers can be shared with torch tensors.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    handle = ctypes.CDLL(str(path))
    for name, restype, argtypes in SIGNATURES:
        fn = getattr(handle, name)
        fn.restype = restype
        fn.argtypes = argtypes
    return handle


def load(require_device: bool = True):
    """Load libsv_vext.so; raise if it is not built or (by default) no GPU is visible."""
    global _lib
    with _lock:
        if _lib is None:
            if not LIB_PATH.exists():
                raise RuntimeError(
                    f"{LIB_PATH} is not built: run `python -m string_vext_amd.build` "
                    "(hipcc, gfx950) before using string_vext_amd")
            _lib = _open(LIB_PATH)
    if require_device and _lib.sv_device_count() < 1:
        raise RuntimeError("libsv_vext.so found no HIP device; the Vext path runs on the GPU only")
    return _lib


def check(code: int) -> None:
    if code:
        msg = _lib.sv_last_error() if _lib is not None else b""
        raise VextError(code, (msg or b"").decode(errors="replace"))
