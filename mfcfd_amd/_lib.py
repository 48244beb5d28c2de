"""ctypes binding of libmfcfd_hip.so (the C ABI of include/mfcfd.h).

The library is loaded from the package directory only; if it is missing the import
fails loudly — there is no CPU fallback for the MI355X path.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmfcfd_hip.so")

MF_OK = 0
MF_ERR_HIP = -1
MF_ERR_INVALID = -2
MF_ERR_STATE = -3
MF_FP64 = 0
MF_MIXED = 1

MF_FIELD_Q = 1
MF_FIELD_DQ = 2
MF_FIELD_QM = 3
MF_FIELD_FLUX_RES = 4
MF_FIELD_DELTA = 5
MF_FIELD_PRIM_OLD = 6
MF_FIELD_NX = 7
MF_FIELD_NY = 8
MF_NUM_KERNEL_KINDS = 5
# mf_adjoint*: snaps value that tapes every iteration in the primal sweep
MF_ADJ_TAPE_ALL = -1
MF_NOALLOC_MAX_SHAPES = 8
# mf_tuning: staged walks (MF_LPP_STAGED | lanes: 1, 2, 4 for LSQ and inner, 2, 4, 8 for flux);
# resident inner sweeps (MF_LPP_RESIDENT | 1, 2 or 4, mixed contexts)
MF_LPP_STAGED = 16
MF_LPP_RESIDENT = 32
# the adjoint's inner reverse in its pull form (MF_LPP_PULL | 1, 2 or 4)
MF_LPP_PULL = 64
KERNEL_KINDS = ("lsq", "inner", "flux_update", "finalize", "forces")


class MfGeometry(C.Structure):
    _fields_ = [("n_points", C.c_int32), ("max_nbhs", C.c_int32),
                ("x", C.c_void_p), ("y", C.c_void_p), ("nx", C.c_void_p), ("ny", C.c_void_p),
                ("min_dist", C.c_void_p), ("left", C.c_void_p), ("right", C.c_void_p),
                ("flag_1", C.c_void_p), ("flag_2", C.c_void_p), ("nbhs", C.c_void_p),
                ("conn", C.c_void_p)]


class MfParams(C.Structure):
    _fields_ = [("cfl", C.c_double), ("vl_const", C.c_double), ("power", C.c_double),
                ("mach", C.c_double), ("aoa", C.c_double), ("euler", C.c_double),
                ("res_old", C.c_double), ("rks", C.c_int32), ("inner_iterations", C.c_int32),
                ("shapes", C.c_int32), ("interior_points_normal_flag", C.c_int32),
                ("format", C.c_int32), ("restart", C.c_int32), ("precision", C.c_int32),
                ("reorder", C.c_int32), ("block_size", C.c_int32), ("first_order", C.c_int32)]


class MfLayout(C.Structure):
    _fields_ = [("n_points", C.c_int64), ("n_wall", C.c_int64), ("n_outer", C.c_int64),
                ("n_interior", C.c_int64), ("n_shape", C.c_int64), ("max_nbhs", C.c_int64),
                ("total_nbhs", C.c_int64), ("state_bytes", C.c_int64)]


class MfAdjointStats(C.Structure):
    _fields_ = [("forward_steps", C.c_int64), ("taped_steps", C.c_int64),
                ("snapshots", C.c_int32), ("reserved", C.c_int32),
                ("primal_ms", C.c_double), ("adjoint_ms", C.c_double),
                ("lpp_flux_rev", C.c_int32), ("lpp_lsq_rev", C.c_int32),
                ("lpp_inner_rev", C.c_int32), ("lpp_gather", C.c_int32)]


class MfCase(C.Structure):
    _fields_ = [("geometry", C.POINTER(MfGeometry)), ("mach", C.c_double), ("aoa", C.c_double)]


class MfTuning(C.Structure):
    _fields_ = [("lpp_lsq", C.c_int32), ("lpp_inner", C.c_int32), ("lpp_flux", C.c_int32),
                ("barrier_polls", C.c_int32), ("us_lsq", C.c_double), ("us_inner", C.c_double),
                ("us_flux", C.c_double), ("lpp_flux_rev", C.c_int32), ("lpp_lsq_rev", C.c_int32),
                ("lpp_inner_rev", C.c_int32), ("lpp_gather", C.c_int32)]


# (name, restype, argtypes) of every symbol include/mfcfd.h declares
_VP, _DP, _IP = C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_int32)
SIGNATURES = [
    ("mf_default_params", None, [C.POINTER(MfParams)]),
    ("mf_create", C.c_int, [C.POINTER(MfGeometry), C.POINTER(MfParams), C.POINTER(C.c_void_p)]),
    ("mf_destroy", C.c_int, [_VP]),
    ("mf_last_error", C.c_char_p, []),
    ("mf_set_prim", C.c_int, [_VP, _DP]),
    ("mf_get_prim", C.c_int, [_VP, _DP]),
    ("mf_get_field", C.c_int, [_VP, C.c_int32, _DP]),
    ("mf_set_field", C.c_int, [_VP, C.c_int32, _DP]),
    ("mf_func_delta", C.c_int, [_VP]),
    ("mf_eval_q_variables", C.c_int, [_VP]),
    ("mf_eval_q_derivatives", C.c_int, [_VP]),
    ("mf_eval_q_inner_loop", C.c_int, [_VP]),
    ("mf_cal_flux_residual", C.c_int, [_VP]),
    ("mf_state_update", C.c_int, [_VP, C.c_int32, _DP]),
    ("mf_compute_cl_cd_cm", C.c_int, [_VP, _DP, _DP, _DP, _DP]),
    ("mf_shape_points", C.c_int32, [_VP]),
    ("mf_run", C.c_int, [_VP, C.c_int32, C.c_int32, _DP, _DP, _DP, _IP]),
    ("mf_run_async", C.c_int, [_VP, C.c_int32, C.c_int32]),
    ("mf_sync", C.c_int, [_VP]),
    ("mf_profile_kernels", C.c_int, [_VP, C.c_int32, _DP, _IP]),
    ("mf_time_kernels", C.c_int, [_VP, C.c_int32, _DP]),
    ("mf_layout", C.c_int, [_VP, C.POINTER(MfLayout)]),
    ("mf_create_batch", C.c_int, [C.POINTER(MfCase), C.c_int32, C.POINTER(MfParams), C.POINTER(C.c_void_p)]),
    ("mf_batch_run", C.c_int, [_VP, C.c_int32, C.c_int32, _DP, _DP, _DP]),
    ("mf_batch_cases", C.c_int, [_VP, _IP, _IP]),
    ("mf_autotune", C.c_int, [_VP, C.c_int32, C.POINTER(MfTuning)]),
    ("mf_get_tuning", C.c_int, [_VP, C.POINTER(MfTuning)]),
    ("mf_set_tuning", C.c_int, [_VP, C.POINTER(MfTuning)]),
    ("mf_tangent", C.c_int, [_VP, C.c_int32, _DP, _DP, _DP, _DP, _DP]),
    ("mf_copy_probe", C.c_int, [C.c_int64, C.c_int32, _DP]),
    # device pointers and the stream travel as plain addresses
    ("mf_compute_cl_cd_cm_noalloc", C.c_int, [C.c_int32, C.c_int32, C.c_int32, C.c_double, C.c_double]
     + [_VP] * 12),
    ("mf_adjoint_forward", C.c_int, [_VP, C.c_int32, C.c_int32, C.c_int32, _DP]),
    ("mf_adjoint_backward", C.c_int, [_VP, _DP, _DP, C.POINTER(MfAdjointStats)]),
    ("mf_adjoint", C.c_int, [_VP, C.c_int32, C.c_int32, C.c_int32, _DP, _DP, _DP,
                             C.POINTER(MfAdjointStats)]),
]

_lib = None


class MfError(RuntimeError):
    pass


def lib():
    """Load libmfcfd_hip.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise MfError(f"{LIB_PATH} is missing: build it with `python -m mfcfd_amd.build`")
        L = C.CDLL(LIB_PATH)
        for name, res, args in SIGNATURES:
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(rc: int, what: str = "") -> None:
    if rc != MF_OK:
        msg = lib().mf_last_error().decode(errors="replace")
        raise MfError(f"{what}: error {rc}: {msg}")
