"""ctypes binding of ``libdiffcp_amd.so`` (the C-ABI in include/diffcp.h).

There is no fallback: if the library is missing the import of this module
raises, naming the build command.
"""
from __future__ import annotations

import ctypes
import os

LIB_PATH = os.environ.get("DIFFCP_LIB") or os.path.join(os.path.dirname(os.path.abspath(__file__)),
                                                        "libdiffcp_amd.so")

c_int = ctypes.c_int32
c_dbl_p = ctypes.POINTER(ctypes.c_double)
c_flt_p = ctypes.POINTER(ctypes.c_float)
c_int_p = ctypes.POINTER(ctypes.c_int32)
c_void_p = ctypes.c_void_p

N_PARAMS = 6
STATUS_NONFINITE = 1
STATUS_OUT_OF_GRID = 2
STATUS_TILE_SPILL = 4


class InstanceDesc(ctypes.Structure):
    _fields_ = [
        ("n_grid", c_int), ("n_vertices", c_int), ("n_elements", c_int), ("n_cube", c_int),
        ("n_grasper", c_int), ("n_grasp_points", c_int), ("n_actions", c_int),
        ("steps_per_frame", c_int), ("frames_per_window", c_int), ("reserved", c_int),
        ("dt", ctypes.c_double), ("gravity", ctypes.c_double),
        ("vertex_p_mass", ctypes.c_double), ("quadrature_p_mass", ctypes.c_double),
        ("quadrature_p_vol", ctypes.c_double), ("cube_p_mass", ctypes.c_double),
        ("operator_p_mass", ctypes.c_double),
        ("vertex_x", c_dbl_p), ("triangles", c_int_p), ("cube_x", c_dbl_p),
        ("grasper_x", c_dbl_p), ("trajectory", c_dbl_p),
    ]


class Options(ctypes.Structure):
    _fields_ = [
        ("max_states", c_int), ("precision", c_int), ("svd_grad", c_int),
        ("per_element_moduli", c_int), ("patch_size", c_int), ("scatter_wgs", c_int),
        ("grid_wgs", c_int), ("bin_cells", c_int),
        ("grid_snapshots", c_int), ("return_mapping", c_int), ("grid_pool_blocks", c_int),
        ("deterministic", c_int),
    ]


# name -> (restype, argtypes); one entry per declaration in include/diffcp.h
SIGNATURES = {
    "dcp_create": (c_int, [ctypes.POINTER(InstanceDesc), c_int, ctypes.POINTER(Options), c_void_p,
                           ctypes.POINTER(c_void_p)]),
    "dcp_destroy": (c_int, [c_void_p]),
    "dcp_last_error": (ctypes.c_char_p, []),
    "dcp_set_params": (c_int, [c_void_p, c_dbl_p]),
    "dcp_set_element_moduli": (c_int, [c_void_p, c_dbl_p]),
    "dcp_set_targets": (c_int, [c_void_p, c_dbl_p, c_int_p]),
    "dcp_reset": (c_int, [c_void_p, c_int, c_dbl_p, c_dbl_p, c_dbl_p]),
    "dcp_rebase_elements": (c_int, [c_void_p, c_int]),
    "dcp_substep": (c_int, [c_void_p, c_int, c_int, c_int, c_int]),
    "dcp_substep_adjoint": (c_int, [c_void_p, c_int, c_int, c_int, c_int, c_int]),
    "dcp_zero_adjoint": (c_int, [c_void_p, c_int]),
    "dcp_zero_param_grads": (c_int, [c_void_p]),
    "dcp_loss": (c_int, [c_void_p, c_int, c_dbl_p]),
    "dcp_loss_adjoint": (c_int, [c_void_p, c_int, c_int, c_dbl_p]),
    "dcp_param_grads": (c_int, [c_void_p, c_dbl_p]),
    "dcp_element_moduli_grads": (c_int, [c_void_p, c_dbl_p]),
    "dcp_set_trajectory": (c_int, [c_void_p, c_dbl_p]),
    "dcp_trajectory_grads": (c_int, [c_void_p, c_dbl_p]),
    "dcp_download_state": (c_int, [c_void_p, c_int] + [c_flt_p] * 11),
    "dcp_upload_state": (c_int, [c_void_p, c_int] + [c_flt_p] * 11),
    "dcp_download_adjoint": (c_int, [c_void_p, c_int] + [c_flt_p] * 11),
    "dcp_upload_adjoint": (c_int, [c_void_p, c_int] + [c_flt_p] * 11),
    "dcp_binning_keys": (c_int, [c_void_p, c_int, c_int_p]),
    "dcp_patch_order": (c_int, [c_void_p, c_int, c_int_p]),
    "dcp_status": (c_int, [c_void_p, ctypes.POINTER(ctypes.c_uint32), c_int]),
    "dcp_rollout_grad": (c_int, [c_void_p, c_int, c_int, c_int, c_int, c_dbl_p, c_dbl_p]),
    "dcp_rollout": (c_int, [c_void_p, c_int, c_int, c_int, c_int_p]),
    "dcp_last_rollout_ms": (c_int, [c_void_p, c_dbl_p]),
    "dcp_sizes": (c_int, [c_void_p, ctypes.POINTER(ctypes.c_int64)]),
    "dcp_device_bytes": (c_int, [c_void_p, ctypes.POINTER(ctypes.c_int64)]),
    "dcp_pool_blocks": (c_int, [c_void_p, c_int_p, c_int_p]),
    "dcp_synchronize": (c_int, [c_void_p]),
    "dcp_polar_adjoint": (c_int, [c_flt_p, c_flt_p, c_int, c_int, c_int, c_flt_p, c_flt_p]),
    "dcp_polar_adjoint_noalloc": (c_int, [c_flt_p, c_flt_p, c_int, c_int, c_int, c_flt_p, c_flt_p]),
    "dcp_last_rollout_final_slot": (c_int, [c_void_p, c_int_p]),
    "dcp_set_kernel_timing": (c_int, [c_void_p, c_int, c_int]),
    "dcp_kernel_timing": (c_int, [c_void_p, c_dbl_p, c_int_p]),
}

_LIB = None


def load() -> ctypes.CDLL:
    """Open the HIP library (once) and declare every C-ABI symbol."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'` "
                          "or `python diffcp_amd/build.py` (requires hipcc)")
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _LIB = lib
    return lib


class DiffcpError(RuntimeError):
    pass


def check(rc: int) -> None:
    if rc != 0:
        msg = load().dcp_last_error()
        raise DiffcpError(msg.decode() if msg else f"diffcp error {rc}")
