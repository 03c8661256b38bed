"""Loader for ``libpvl_hip.so`` (the gfx950 HIP library) and its C-ABI signatures.

There is no fallback: if the library is missing or cannot be loaded, every entry point
raises.  The signatures mirror ``include/pvl.h``.
"""

from __future__ import annotations

import atexit
import ctypes as C
import os
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "libpvl_hip.so"

_u32p = C.POINTER(C.c_uint32)
_uintp = C.POINTER(C.c_uint)
_dp = C.c_void_p  # device pointer
_hp = C.c_void_p  # host pointer

PVL_OK = 0
PVL_ERR_ARG = -1
PVL_ERR_HIP = -2
PVL_ERR_NOMEM = -3
PVL_ERR_SINGULAR = -4
PVL_ERR_NOTCONV = -5
PVL_ERR_NODEV = -6

PVL_FP64 = 0
PVL_FP32 = 1
PVL_FP64_REF = 2

# name -> (restype, argtypes)
SIGNATURES: dict[str, tuple] = {
    "pvl_last_error": (C.c_char_p, []),
    "pvl_init": (C.c_int, [C.c_int]),
    "pvl_device_count": (C.c_int, [C.POINTER(C.c_int)]),
    "pvl_synchronize": (C.c_int, []),
    "pvl_shutdown": (C.c_int, []),
    "pvl_malloc": (C.c_int, [C.POINTER(C.c_void_p), C.c_size_t]),
    "pvl_free": (C.c_int, [_dp]),
    "pvl_memcpy_h2d": (C.c_int, [_dp, _hp, C.c_size_t]),
    "pvl_memcpy_d2h": (C.c_int, [_hp, _dp, C.c_size_t]),
    "pvl_memcpy_d2d": (C.c_int, [_dp, _dp, C.c_size_t]),
    "pvl_memset_zero": (C.c_int, [_dp, C.c_size_t]),
    "pvl_prof_enable": (C.c_int, [C.c_int]),
    "pvl_prof_filter": (C.c_int, [C.c_char_p]),
    "pvl_prof_reset": (C.c_int, []),
    "pvl_prof_query": (C.c_int, [C.c_char_p, C.POINTER(C.c_double), C.POINTER(C.c_long)]),
    "pvl_prof_work": (C.c_int, [C.c_char_p, C.POINTER(C.c_double)]),
    "pvl_timer_start": (C.c_int, []),
    "pvl_timer_stop": (C.c_int, [C.POINTER(C.c_double)]),
    "pvl_mesh_from_elements": (
        C.c_int,
        [C.c_uint, _uintp, _uintp, _uintp, _u32p, _u32p, _u32p],
    ),
    "pvl_mesh_dual": (
        C.c_int,
        [C.c_uint, C.c_uint, _u32p, C.c_uint, _u32p, _u32p, _u32p, _u32p, _u32p],
    ),
    "pvl_wake_topology": (C.c_int, [C.c_uint, C.c_uint, _uintp, _u32p, _u32p, _u32p]),
    "pvl_mesh_create": (
        C.c_int,
        [C.POINTER(C.c_void_p), C.c_uint, C.c_uint, _u32p, C.c_uint, _u32p, _u32p],
    ),
    "pvl_mesh_destroy": (C.c_int, [C.c_void_p]),
    "pvl_mesh_counts": (C.c_int, [C.c_void_p, _uintp, _uintp, _uintp]),
    "pvl_induction_matrix3": (
        C.c_int,
        [C.c_void_p, C.c_double, _dp, C.c_uint, _dp, _dp, _dp, C.c_int],
    ),
    "pvl_induction_matrix": (C.c_int, [C.c_void_p, C.c_double, _dp, C.c_uint, _dp, _dp, C.c_int]),
    "pvl_line_induction_matrix": (
        C.c_int,
        [C.c_void_p, C.c_double, _dp, C.c_uint, _dp, _dp, C.c_int],
    ),
    "pvl_point_velocity": (
        C.c_int,
        [C.c_void_p, C.c_double, _dp, _dp, C.c_uint, _dp, _dp, C.c_int, C.c_int],
    ),
    "pvl_point_velocity_noalloc": (
        C.c_int,
        [C.c_void_p, C.c_double, _dp, _dp, C.c_uint, _dp, _dp, C.c_int, C.c_int],
    ),
    "pvl_line_forces": (C.c_int, [C.c_void_p, _dp, C.c_uint, C.c_uint, C.c_uint, _dp, _dp, _dp, _dp]),
    "pvl_surface_normal": (C.c_int, [C.c_void_p, _dp, _dp]),
    "pvl_surface_center": (C.c_int, [C.c_void_p, _dp, _dp]),
    "pvl_induction_matrix3_host": (
        C.c_int,
        [C.c_void_p, C.c_double, _hp, C.c_uint, _hp, _hp, _hp, C.c_int],
    ),
    "pvl_induction_matrix_host": (
        C.c_int,
        [C.c_void_p, C.c_double, _hp, C.c_uint, _hp, _hp, C.c_int],
    ),
    "pvl_line_induction_matrix_host": (
        C.c_int,
        [C.c_void_p, C.c_double, _hp, C.c_uint, _hp, _hp, C.c_int],
    ),
    "pvl_lu_create": (C.c_int, [C.POINTER(C.c_void_p), C.c_int, C.c_int]),
    "pvl_lu_set_pivoting": (C.c_int, [C.c_void_p, C.c_int]),
    "pvl_lu_destroy": (C.c_int, [C.c_void_p]),
    "pvl_lu_factor": (C.c_int, [C.c_void_p, _dp, C.POINTER(C.c_int)]),
    "pvl_lu_solve": (C.c_int, [C.c_void_p, _dp, C.c_int]),
    "pvl_lu_pivots": (C.c_int, [C.c_void_p, C.POINTER(C.c_int)]),
    "pvl_lu_factors": (C.c_int, [C.c_void_p, _dp]),
    "pvl_residual": (C.c_int, [C.c_int, _dp, _dp, _dp, _dp, C.c_int]),
    "pvl_solve_refine": (
        C.c_int,
        [C.c_void_p, _dp, _dp, _dp, C.c_int, C.c_double, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_double)],
    ),
    "pvl_deflate": (C.c_int, [C.c_int, _dp, C.c_int, C.c_int, C.c_double]),
    "pvl_wake_create": (
        C.c_int,
        [C.POINTER(C.c_void_p), C.c_uint, C.c_uint, _u32p, _u32p, C.c_double, C.c_double],
    ),
    "pvl_wake_destroy": (C.c_int, [C.c_void_p]),
    "pvl_wake_info": (C.c_int, [C.c_void_p, _uintp, _uintp, _uintp, C.POINTER(C.c_double)]),
    "pvl_wake_update": (
        C.c_int,
        [C.c_void_p, C.c_double, C.c_void_p, _dp, _dp, C.POINTER(C.c_double), C.c_int, C.c_int],
    ),
    "pvl_wake_update_flow": (
        C.c_int,
        [C.c_void_p, C.c_double, C.c_void_p, _dp, _dp, C.POINTER(C.c_double), C.c_int, C.c_int],
    ),
    "pvl_wake_begin_update": (C.c_int, [C.c_void_p, C.c_int, C.c_int]),
    "pvl_wake_apply_corrections": (
        C.c_int,
        [C.c_void_p, C.c_uint, _dp, _dp, _dp, C.c_uint, _dp, C.c_int],
    ),
    "pvl_wake_get_velocity": (C.c_int, [C.c_void_p, C.c_uint, _dp, _dp, C.c_int, C.c_int]),
    "pvl_wake_get_state": (C.c_int, [C.c_void_p, _hp, _hp]),
    "pvl_wake_set_state": (C.c_int, [C.c_void_p, _hp, _hp, C.c_uint, C.c_double]),
    "pvl_mlp_create": (C.c_int, [C.POINTER(C.c_void_p), C.c_int, C.c_int, C.c_int]),
    "pvl_mlp_destroy": (C.c_int, [C.c_void_p]),
    "pvl_mlp_num_params": (C.c_int, [C.c_void_p, C.POINTER(C.c_long)]),
    "pvl_mlp_set_params": (C.c_int, [C.c_void_p, _hp]),
    "pvl_mlp_get_params": (C.c_int, [C.c_void_p, _hp]),
    "pvl_mlp_forward": (C.c_int, [C.c_void_p, _dp, _dp, C.c_int]),
    "pvl_mlp_gradient": (C.c_int, [C.c_void_p, _dp, _dp, C.c_int, _hp, C.POINTER(C.c_float)]),
    "pvl_mlp_train_step": (
        C.c_int,
        [C.c_void_p, _dp, _dp, C.c_int, C.c_float, C.c_float, C.c_float, C.c_float, C.c_float,
         C.POINTER(C.c_float)],
    ),
    "pvl_gmres_refine": (
        C.c_int,
        [C.c_int, _dp, _dp, _dp, _dp, C.c_double, C.c_int, C.c_double, C.c_int, C.POINTER(C.c_int),
         C.POINTER(C.c_double)],
    ),
    "pvl_convert_f64_f32": (C.c_int, [_dp, _dp, C.c_size_t]),
    "pvl_rhs_normal_velocity": (C.c_int, [C.c_uint, _dp, _dp, C.POINTER(C.c_double), _dp]),
    "pvl_rhs_flow": (C.c_int, [C.c_uint, _dp, _dp, _dp, C.POINTER(C.c_double), _dp]),
    "pvl_remove_mean": (C.c_int, [_dp, C.c_int, C.c_int]),
    "pvl_frame_kinematics": (
        C.c_int,
        [C.c_uint, _dp, C.c_uint, C.POINTER(C.c_double), C.POINTER(C.c_int), _dp, _dp],
    ),
}

_LIB: C.CDLL | None = None
_SHUT = False


class PvlError(RuntimeError):
    """A libpvl_hip.so call returned a failure status."""

    def __init__(self, status: int, message: str):
        super().__init__(message)
        self.status = status


def load() -> C.CDLL:
    """Load the HIP library (once) and attach the C-ABI signatures."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not LIB_PATH.exists():
        raise RuntimeError(
            f"{LIB_PATH} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'` "
            "(there is no CPU fallback)"
        )
    lib = C.CDLL(os.fspath(LIB_PATH), mode=C.RTLD_GLOBAL)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _LIB = lib
    atexit.register(_shutdown)
    return lib


def _shutdown() -> None:
    """Release the library's streams, events, pinned memory and workspaces while the HIP
    runtime is still up (atexit runs before interpreter and runtime teardown).  Handles
    still alive afterwards are left to process exit: their finalizers become no-ops."""
    global _SHUT
    if _LIB is not None and not _SHUT:
        _SHUT = True
        _LIB.pvl_shutdown()


def alive() -> bool:
    """True while handles may still be released through the library."""
    return _LIB is not None and not _SHUT


def check(status: int, what: str = "") -> None:
    """Raise the Python exception matching a libpvl status code."""
    if status == PVL_OK:
        return
    msg = load().pvl_last_error().decode(errors="replace")
    if what:
        msg = f"{what}: {msg}"
    if status == PVL_ERR_ARG:
        raise ValueError(msg)
    if status == PVL_ERR_NOMEM:
        raise MemoryError(msg)
    raise PvlError(status, msg)


def call(name: str, *args) -> None:
    """Call a status-returning entry point and raise on failure."""
    check(getattr(load(), name)(*args), name)
