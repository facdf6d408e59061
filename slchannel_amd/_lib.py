"""ctypes binding of libslc_hip.so (the C ABI declared in include/slc.h,
include/slc_ens.h and include/slc_stats.h).

The library is loaded from the package directory only; if it is missing the
import of any GPU entry point raises -- there is no CPU fallback.
"""

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
# SLCHANNEL_LIB overrides the library file (e.g. a build with another tile shape)
LIB_PATH = os.environ.get("SLCHANNEL_LIB") or os.path.join(HERE, "libslc_hip.so")

# Every symbol include/slc.h, include/slc_ens.h and include/slc_stats.h
# declare (checked by the CPU test suite).
EXPORTS = (
    "slc_version", "slc_last_error", "slc_plan_create", "slc_plan_destroy",
    "slc_plan_workspace_bytes", "slc_apply_bc", "slc_departure", "slc_gather",
    "slc_stencil", "slc_locate_z", "slc_advect", "slc_bdf1_predict", "slc_bdf2_predict",
    "slc_zsolve", "slc_divergence", "slc_poisson", "slc_project", "slc_project_step",
    "slc_pressure_combine",
    "slc_bulk_velocity", "slc_bulk_force", "slc_bulk_force_noalloc", "slc_diagnostics",
    "slc_init_solenoidal", "slc_init_solenoidal_work", "slc_interp_trilinear",
    "slc_imex_rhs",
    "slc_ab2_update", "slc_set_option", "slc_set_viscosity",
    "slc_profile_enable", "slc_profile_report",
    "slc_ens_create", "slc_ens_destroy", "slc_ens_member_bytes", "slc_ens_add",
    "slc_ens_remove", "slc_ens_count", "slc_ens_bytes", "slc_ens_trim_pool",
    "slc_ens_set_fields",
    "slc_ens_get_fields", "slc_ens_stats", "slc_ens_step", "slc_ens_set_option",
    "slc_ens_get_option", "slc_ens_autotune", "slc_ens_diagnostics",
    "slc_stats_create", "slc_stats_destroy", "slc_stats_workspace_bytes",
    "slc_stats_accumulate",
)


class GridDesc(ctypes.Structure):
    """slc_grid_desc (include/slc.h)."""

    _fields_ = [
        ("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int),
        ("Lx", ctypes.c_double), ("Ly", ctypes.c_double), ("Lz", ctypes.c_double),
        ("gamma", ctypes.c_double), ("tanh_g", ctypes.c_double),
        ("stretching", ctypes.c_int), ("order", ctypes.c_int), ("traj_order", ctypes.c_int),
        ("n_traj_iters", ctypes.c_int), ("top_wall", ctypes.c_int),
        ("nu", ctypes.c_double),
        ("z_f", ctypes.POINTER(ctypes.c_double)), ("z_c", ctypes.POINTER(ctypes.c_double)),
    ]


_lib = None


def load():
    """Load libslc_hip.so once; raise RuntimeError if it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"libslc_hip.so not found at {LIB_PATH}: build it first "
            "(python -m slchannel_amd.build or __graft_entry__.build())"
        )
    lib = ctypes.CDLL(LIB_PATH)
    vp = ctypes.c_void_p
    i, d, i64p = ctypes.c_int, ctypes.c_double, ctypes.c_void_p
    sig = {
        "slc_version": ([], i),
        "slc_last_error": ([], ctypes.c_char_p),
        "slc_plan_create": ([ctypes.POINTER(GridDesc), i, ctypes.POINTER(vp)], i),
        "slc_plan_destroy": ([vp], i),
        "slc_plan_workspace_bytes": ([vp, ctypes.POINTER(ctypes.c_int64)], i),
        "slc_apply_bc": ([vp, vp, vp, vp, vp], i),
        "slc_departure": ([vp, i, vp, vp, vp, i, d, vp, vp, vp, i64p, vp], i),
        "slc_gather": ([vp, i, vp, i, vp, vp, vp, vp, i64p, vp], i),
        "slc_stencil": ([vp, i, vp, vp, vp, vp, vp, vp, i64p, vp], i),
        "slc_locate_z": ([vp, i, vp, ctypes.c_int64, vp, vp, vp], i),
        "slc_advect": ([vp, vp, vp, vp, vp, vp, vp, d, vp, vp, vp, i64p, vp], i),
        "slc_bdf1_predict": ([vp, vp, vp, vp, d, vp, vp, vp, i64p, vp], i),
        "slc_bdf2_predict": ([vp, vp, vp, vp, vp, vp, vp, d, vp, vp, vp, i64p, vp], i),
        "slc_zsolve": ([vp, i, vp, d, d, vp], i),
        "slc_divergence": ([vp, vp, vp, vp, vp, vp], i),
        "slc_poisson": ([vp, vp, d, vp, vp], i),
        "slc_project": ([vp, vp, vp, vp, vp, d, vp], i),
        "slc_pressure_combine": ([vp, vp, d, vp, d, vp, vp], i),
        "slc_project_step": ([vp, vp, vp, vp, d, d, vp, vp], i),
        "slc_bulk_velocity": ([vp, vp, vp, vp], i),
        "slc_bulk_force": ([vp, vp, vp, vp, d, d, vp, vp], i),
        "slc_bulk_force_noalloc": ([vp, vp, vp, vp, d, d, vp, vp], i),
        "slc_diagnostics": ([vp, vp, vp, vp, vp, d, vp, vp], i),
        "slc_init_solenoidal": ([vp, i, i, vp, vp, vp, d, vp, ctypes.c_int64, vp, vp, vp, vp], i),
        "slc_init_solenoidal_work": ([i, i, i, i], ctypes.c_int64),
        "slc_interp_trilinear": ([i, vp, i, i, i, vp, vp, i, vp, vp, i, vp, vp, i, vp, vp], i),
        "slc_imex_rhs": ([vp, vp, vp, vp, vp, vp, vp, vp], i),
        "slc_ab2_update": ([vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, d, vp], i),
        "slc_set_option": ([vp, i, i], i),
        "slc_set_viscosity": ([vp, d], i),
        "slc_profile_enable": ([vp, i], i),
        "slc_profile_report": ([vp, ctypes.c_char_p, ctypes.c_int64], i),
        "slc_ens_create": ([i, i, i, ctypes.POINTER(vp)], i),
        "slc_ens_destroy": ([vp], i),
        "slc_ens_member_bytes": ([ctypes.POINTER(GridDesc), i, ctypes.POINTER(ctypes.c_int64)], i),
        "slc_ens_add": ([vp, ctypes.POINTER(GridDesc), d, d, ctypes.POINTER(i)], i),
        "slc_ens_remove": ([vp, i], i),
        "slc_ens_count": ([vp, ctypes.POINTER(i)], i),
        "slc_ens_bytes": ([vp, ctypes.POINTER(ctypes.c_int64)], i),
        "slc_ens_trim_pool": ([vp, ctypes.c_int64, ctypes.POINTER(ctypes.c_int64)], i),
        "slc_ens_set_fields": ([vp, i, vp, vp, vp, vp, vp], i),
        "slc_ens_get_fields": ([vp, i, vp, vp, vp, vp, vp], i),
        "slc_ens_stats": ([vp, i, vp, vp], i),
        "slc_ens_step": ([vp, vp], i),
        "slc_ens_set_option": ([vp, i, i], i),
        "slc_ens_get_option": ([vp, i, ctypes.POINTER(i)], i),
        "slc_ens_autotune": ([vp, i, vp, ctypes.c_char_p, ctypes.c_int64], i),
        "slc_ens_diagnostics": ([vp, i, vp, vp], i),
        "slc_stats_create": ([i, i, i, i, i, vp, vp, ctypes.POINTER(vp)], i),
        "slc_stats_destroy": ([vp], i),
        "slc_stats_workspace_bytes": ([vp, ctypes.POINTER(ctypes.c_int64)], i),
        "slc_stats_accumulate": ([vp, vp, vp, vp, vp, vp, vp], i),
    }
    for name, (args, res) in sig.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = res
    _lib = lib
    return lib


def check(rc, what):
    """Raise RuntimeError with the library's message on a non-zero status."""
    if rc != 0:
        msg = load().slc_last_error().decode(errors="replace")
        raise RuntimeError(f"{what} failed (status {rc}): {msg}")
