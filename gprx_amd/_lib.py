"""ctypes binding of libgprx_hip.so (the C ABI in include/gprx_hip.h).

The library is built in-tree by gprx_amd/build.py.  Loading fails loudly when
it is missing: there is no CPU fallback behind this package.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "libgprx_hip.so"

_lib = None

# gprx_status (include/gprx_hip.h)
OK = 0
CHOLESKY_FAILED = 1
NON_FINITE_KERNEL_VALUE = 2
NON_FINITE_INPUT = 3
INVALID_HYPERPARAMETER = 4
INVALID_NOISE_VARIANCE = 5
LENGTH_MISMATCH = 6
EMPTY_INPUT = 7
UNSUPPORTED = 8
DEVICE_ERROR = 9
NOT_CONVERGED = 10
INVALID_CONFIG = 11
INVALID_INTERVAL = 12
INSUFFICIENT_DATA = 13

FP64 = 0
MIXED = 1

VARIANCE_OBSERVATION = 0
VARIANCE_LATENT = 1

FALLBACK_NONE = 0
FALLBACK_REFINEMENT = 1
FALLBACK_JITTER = 2

# Every symbol the header declares, with (restype, argtypes).
_dp = C.POINTER(C.c_double)
_vp = C.c_void_p


class GprxError(C.Structure):
    _fields_ = [
        ("status", C.c_int32),
        ("stage", C.c_int32),
        ("jitter", C.c_double),
        ("matrix_size", C.c_size_t),
        ("message", C.c_char * 256),
    ]


class KernelProgram(C.Structure):
    _fields_ = [("ops", C.POINTER(C.c_int32)), ("n_ops", C.c_size_t)]


class JitterPolicyC(C.Structure):
    _fields_ = [
        ("initial", C.c_double),
        ("multiplier", C.c_double),
        ("max_jitter", C.c_double),
        ("max_retries", C.c_int32),
    ]


class ProblemC(C.Structure):
    _fields_ = [
        ("x", _dp),
        ("n", C.c_size_t),
        ("d", C.c_size_t),
        ("y", _dp),
        ("kernel", C.POINTER(KernelProgram)),
        ("log_theta", _dp),
        ("p", C.c_size_t),
    ]


class LbfgsOptionsC(C.Structure):
    _fields_ = [
        ("max_iterations", C.c_uint64),
        ("tolerance", C.c_double),
        ("history_size", C.c_int32),
        ("restarts", C.c_uint32),
        ("seed", C.c_uint64),
    ]


# gprx_objective_fn: int (*)(void* ctx, const double* x, size_t n, double* value, double* grad)
ObjectiveFn = C.CFUNCTYPE(C.c_int, C.c_void_p, _dp, C.c_size_t, _dp, _dp)
# gprx_objective2_fn: the same with the Hessian (double* hess) last
Objective2Fn = C.CFUNCTYPE(C.c_int, C.c_void_p, _dp, C.c_size_t, _dp, _dp, _dp)


class TrOptionsC(C.Structure):
    _fields_ = [
        ("max_iterations", C.c_uint64),
        ("tolerance", C.c_double),
        ("initial_radius", C.c_double),
        ("max_radius", C.c_double),
        ("restarts", C.c_uint32),
        ("seed", C.c_uint64),
    ]

_err = C.POINTER(GprxError)
_mp = C.POINTER(C.c_void_p)

SIGNATURES = {
    "gprx_hip_version": (C.c_char_p, []),
    "gprx_hip_device_count": (C.c_int, []),
    "gprx_hip_exact_create": (C.c_int, [_dp, C.c_size_t, C.c_size_t, _dp, C.POINTER(KernelProgram),
                                        _dp, C.c_size_t, C.c_int32, C.POINTER(JitterPolicyC), _mp,
                                        _err]),
    "gprx_hip_exact_value_and_grad": (C.c_int, [_vp, _dp, C.c_size_t, _dp, _dp, _err]),
    "gprx_hip_exact_hessian": (C.c_int, [_vp, _dp, C.c_size_t, _dp, _dp, _dp, _err]),
    "gprx_hip_exact_delete": (C.c_int, [_vp, C.c_size_t, _dp, _err]),
    "gprx_hip_exact_export_factor": (C.c_int, [_vp, _dp, C.c_size_t, _dp, _dp, _err]),
    "gprx_hip_exact_insert": (C.c_int, [_vp, _dp, C.c_size_t, C.c_double, _dp, _err]),
    "gprx_hip_exact_predict": (C.c_int, [_vp, _dp, C.c_size_t, C.c_size_t, C.c_int32, _dp, _dp,
                                         _err]),
    "gprx_hip_exact_predict_noalloc": (C.c_int, [_vp, _dp, C.c_size_t, C.c_size_t, C.c_int32, _dp,
                                                 _dp, _err]),
    "gprx_hip_get_params": (C.c_int, [_vp, _dp, C.c_size_t]),
    "gprx_hip_last_jitter": (C.c_double, [_vp]),
    "gprx_hip_last_fallback": (C.c_int, [_vp]),
    "gprx_hip_last_refine_steps": (C.c_int, [_vp]),
    "gprx_hip_set_mixed_jitter_bound": (C.c_int, [_vp, C.c_double, _err]),
    "gprx_hip_sgpr_create": (C.c_int, [_dp, C.c_size_t, C.c_size_t, _dp, _dp, C.c_size_t,
                                       C.POINTER(KernelProgram), _dp, C.c_size_t, C.c_int32, _mp,
                                       _err]),
    "gprx_hip_sgpr_value_and_grad": (C.c_int, [_vp, _dp, C.c_size_t, _dp, _dp, _err]),
    "gprx_hip_sgpr_predict": (C.c_int, [_vp, _dp, C.c_size_t, C.c_size_t, C.c_int32, _dp, _dp,
                                        _err]),
    "gprx_hip_svgp_create": (C.c_int, [_dp, C.c_size_t, C.c_size_t, _dp, _dp, C.c_size_t,
                                       C.POINTER(KernelProgram), _dp, C.c_size_t, C.c_int32, _mp,
                                       _err]),
    "gprx_hip_svgp_num_params": (C.c_size_t, [_vp]),
    "gprx_hip_svgp_value_and_grad": (C.c_int, [_vp, _dp, C.c_size_t, C.POINTER(C.c_uint64),
                                               C.c_size_t, _dp, _dp, _err]),
    "gprx_hip_svgp_adam": (C.c_int, [_vp, _dp, C.c_size_t, C.c_size_t, C.c_size_t, C.c_double,
                                     C.c_double, C.c_double, C.c_double, C.c_uint64, _dp, _err]),
    "gprx_hip_svgp_predict": (C.c_int, [_vp, _dp, C.c_size_t, _dp, C.c_size_t, C.c_size_t,
                                        C.c_int32, _dp, _dp, _err]),
    "gprx_hip_batch_create": (C.c_int, [C.POINTER(ProblemC), C.c_size_t, C.c_int32, _mp, _err]),
    "gprx_hip_batch_value_and_grad": (C.c_int, [_vp, C.POINTER(_dp), _dp, C.POINTER(_dp), _err]),
    "gprx_hip_batch_set_schedule": (C.c_int, [_vp, C.c_int32, C.c_int32, _err]),
    "gprx_hip_batch_schedule": (C.c_int, [_vp, C.POINTER(C.c_int32), C.POINTER(C.c_int32),
                                          C.POINTER(C.c_double)]),
    "gprx_hip_batch_last_fallback": (C.c_int, [_vp, C.POINTER(C.c_int32), C.c_size_t]),
    "gprx_lbfgs_default_options": (None, [C.POINTER(LbfgsOptionsC)]),
    "gprx_lbfgs_minimize": (C.c_int, [ObjectiveFn, _vp, _dp, C.c_size_t, C.POINTER(LbfgsOptionsC),
                                      _dp, _dp, C.POINTER(C.c_uint64), _err]),
    "gprx_hip_exact_fit": (C.c_int, [_dp, C.c_size_t, C.c_size_t, _dp, C.POINTER(KernelProgram), _dp,
                                     C.c_size_t, _dp, C.c_int32, C.POINTER(JitterPolicyC),
                                     C.POINTER(LbfgsOptionsC), _dp, _dp, C.POINTER(C.c_uint64),
                                     C.POINTER(C.c_int32), _err]),
    "gprx_tr_default_options": (None, [C.POINTER(TrOptionsC)]),
    "gprx_trust_region_minimize": (C.c_int, [Objective2Fn, _vp, _dp, C.c_size_t, C.POINTER(TrOptionsC),
                                             _dp, _dp, C.POINTER(C.c_uint64), _err]),
    "gprx_hip_exact_fit_trust_region": (C.c_int, [_vp, _dp, C.c_size_t, _dp, C.POINTER(TrOptionsC), _dp,
                                                  _dp, C.POINTER(C.c_uint64), C.POINTER(C.c_int32),
                                                  _err]),
    "gprx_hip_destroy": (None, [_vp]),
    "gprx_hip_last_eval_ms": (C.c_double, [_vp]),
    "gprx_hip_last_dense_ms": (C.c_double, [_vp]),
    "gprx_hip_last_phase_ms": (C.c_double, [_vp, C.c_int32]),
}


def lib():
    """The loaded library; raises if it was not built."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it with `python gprx_amd/build.py` "
                "(gfx950 HIP library; there is no CPU fallback)")
        L = C.CDLL(str(LIB_PATH))
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def dptr(a: np.ndarray):
    return a.ctypes.data_as(_dp)


def f64(a) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a, dtype=np.float64).reshape(-1))
