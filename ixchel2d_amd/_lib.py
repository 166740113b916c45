"""ctypes binding of libixchel2d.so (the C-ABI declared in include/ixchel2d.h).

The library is built in-tree by ``__graft_entry__.build()``; importing the
binding raises when it is missing -- there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libixchel2d.so")

IXH_U, IXH_V, IXH_T, IXH_P, IXH_U_ANT, IXH_V_ANT, IXH_T_ANT, IXH_PCORR, IXH_BO, IXH_AU, IXH_AV, IXH_RESU = range(12)
FIELD_ID = {"u": IXH_U, "v": IXH_V, "t": IXH_T, "p": IXH_P, "u_ant": IXH_U_ANT, "v_ant": IXH_V_ANT,
            "t_ant": IXH_T_ANT, "pcorr": IXH_PCORR, "bo": IXH_BO, "au": IXH_AU, "av": IXH_AV,
            "resu": IXH_RESU}
VAR_ID = {"u": 0, "v": 1, "t": 2}
SIDE_ID = {"a": 0, "b": 1, "c": 2, "d": 3}
DIR_Y, DIR_X = 0, 1
SOLVER_ID = {"pcr": 0, "thomas": 1, "pcr_exact": 2}
PRECISION_ID = {"fp64": 0, "fp32_mixed": 1, "fp32_refined": 2, "fp32_stochastic": 3}

# every symbol include/ixchel2d.h declares
EXPORTS = [
    "ixh_last_error", "ixh_version", "ixh_create", "ixh_destroy", "ixh_set_params", "ixh_set_stream",
    "ixh_synchronize", "ixh_upload_field", "ixh_download_field", "ixh_set_bc", "ixh_set_sources",
    "ixh_momentum_u", "ixh_momentum_v", "ixh_pcorr_zero", "ixh_pcorr", "ixh_correct", "ixh_energy",
    "ixh_residual_u", "ixh_lag", "ixh_simple_iterations", "ixh_nusselt_y", "ixh_nusselt_x",
    "ixh_tridiagonal", "ixh_tridiagonal_noalloc", "ixh_set_profiling", "ixh_kernel_times", "ixh_kernel_name",
    "ixh_batch_create", "ixh_batch_destroy", "ixh_batch_set_schedule", "ixh_batch_set_shape",
    "ixh_batch_iterations",
    "ixh_batch_autotune", "ixh_adjoint_record", "ixh_adjoint_energy",
    "ixh_adjoint_coupled",
]
ORDER_ID = {"cavity": 0, "largest": 1, "xcd": 2}
BATCH_SCHEDULES = 48  # entries of ixh_batch_autotune's timing table
BATCH_LINE_GROUP = 4  # lines per workgroup of the grouped batch sweeps


class Grid(C.Structure):
    _fields_ = [("mi", C.c_int32), ("nj", C.c_int32), ("xu", C.POINTER(C.c_double)),
                ("yv", C.POINTER(C.c_double)), ("xp", C.POINTER(C.c_double)),
                ("yp", C.POINTER(C.c_double))]


class Params(C.Structure):
    _fields_ = [("dt", C.c_double), ("rel_pres", C.c_double), ("rel_vel", C.c_double),
                ("rel_ener", C.c_double), ("gamma_momen", C.c_double), ("gamma_energ", C.c_double),
                ("ri_x", C.c_double), ("ri_y", C.c_double), ("conv_u", C.c_double),
                ("conv_t", C.c_double), ("conv_p", C.c_double), ("conv_resi", C.c_double),
                ("conv_paso", C.c_double), ("iter_ecuaci_max", C.c_int32),
                ("iter_simple_max", C.c_int32), ("solver", C.c_int32), ("precision", C.c_int32)]


class Hist(C.Structure):
    _fields_ = [("err_v", C.c_double), ("err_p", C.c_double), ("maxbo", C.c_double),
                ("err_t", C.c_double), ("residuo", C.c_double), ("passes_mom", C.c_int32),
                ("passes_p", C.c_int32), ("passes_t", C.c_int32), ("simple_iter", C.c_int32),
                ("logged", C.c_int32), ("pad_", C.c_int32)]


class IxhError(RuntimeError):
    pass


_lib = None


def load(path: str = LIB_PATH):
    """Load and prototype the library (no GPU call is made)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise IxhError(f"{path} is not built; run __graft_entry__.build()")
    L = C.CDLL(path)
    dp = C.POINTER(C.c_double)
    ip = C.POINTER(C.c_int32)
    vp = C.c_void_p
    L.ixh_last_error.restype = C.c_char_p
    L.ixh_version.restype = C.c_char_p
    L.ixh_kernel_name.restype = C.c_char_p
    L.ixh_kernel_name.argtypes = [C.c_int32]
    L.ixh_create.argtypes = [C.POINTER(Grid), C.POINTER(Params), C.POINTER(vp)]
    L.ixh_destroy.argtypes = [vp]
    L.ixh_set_params.argtypes = [vp, C.POINTER(Params)]
    L.ixh_set_stream.argtypes = [vp, vp]
    L.ixh_synchronize.argtypes = [vp]
    L.ixh_upload_field.argtypes = [vp, C.c_int32, dp, C.c_int32]
    L.ixh_download_field.argtypes = [vp, C.c_int32, dp, C.c_int32]
    L.ixh_set_bc.argtypes = [vp, C.c_int32, C.c_int32, ip, dp]
    L.ixh_set_sources.argtypes = [vp, C.c_int32, dp, dp]
    L.ixh_momentum_u.argtypes = [vp, C.c_int32]
    L.ixh_momentum_v.argtypes = [vp, C.c_int32, dp]
    L.ixh_pcorr_zero.argtypes = [vp]
    L.ixh_pcorr.argtypes = [vp, C.c_int32, dp, dp]
    L.ixh_correct.argtypes = [vp]
    L.ixh_energy.argtypes = [vp, C.c_int32, dp]
    L.ixh_residual_u.argtypes = [vp, dp]
    L.ixh_lag.argtypes = [vp]
    L.ixh_simple_iterations.argtypes = [vp, C.c_int32, C.POINTER(Hist), C.c_int32, ip]
    L.ixh_nusselt_y.argtypes = [vp, C.c_int32, C.c_int32, dp, dp]
    L.ixh_nusselt_x.argtypes = [vp, dp, dp]
    L.ixh_tridiagonal.argtypes = [dp, dp, dp, dp, C.c_int32, C.c_int32, C.c_int32]
    L.ixh_tridiagonal_noalloc.argtypes = [dp, dp, dp, dp, C.c_int32, C.c_int32, C.c_int32]
    L.ixh_set_profiling.argtypes = [vp, C.c_int32]
    L.ixh_kernel_times.argtypes = [vp, dp, ip, C.c_int32, ip]
    L.ixh_batch_create.argtypes = [C.POINTER(vp), C.c_int32, C.POINTER(vp)]
    L.ixh_batch_destroy.argtypes = [vp]
    L.ixh_batch_set_schedule.argtypes = [vp, C.c_int32, C.c_int32]
    L.ixh_batch_iterations.argtypes = [vp, C.c_int32, C.POINTER(Hist), C.c_int32]
    L.ixh_batch_set_shape.argtypes = [vp, C.c_int32, C.c_int32, C.c_int32]
    L.ixh_batch_autotune.argtypes = [vp, C.c_int32, ip, ip, ip, ip, ip, dp]
    L.ixh_adjoint_record.argtypes = [vp, C.c_int32, C.c_int32]
    L.ixh_adjoint_energy.argtypes = [vp, dp, C.c_int32, dp, dp, dp, dp, dp, C.c_int32]
    L.ixh_adjoint_coupled.argtypes = [vp, dp, C.c_int32, dp, dp, dp, dp, dp, C.c_int32]
    for name in EXPORTS:
        fn = getattr(L, name)
        if fn.restype is C.c_int or fn.restype is None:
            if name not in ("ixh_last_error", "ixh_version", "ixh_kernel_name"):
                fn.restype = C.c_int
    _lib = L
    return L


def check(rc: int) -> None:
    if rc != 0:
        msg = load().ixh_last_error().decode(errors="replace")
        raise IxhError(f"ixchel2d error {rc}: {msg}")
