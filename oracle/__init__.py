"""CPU f64 oracle of gprx's exact / VFE / SVGP hot path (test infrastructure).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load
this package, and only as the checker or the timed CPU baseline.  The product
(gprx_amd) never calls it.

Parity pinning: see the header of gprx_oracle.c.  Every function here is
checked in tests/test_oracle_*.py against the reference's golden fixtures,
closed forms, finite differences or invariants.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "liboracle.so"

_lib = None


def build() -> Path:
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    return LIB


def lib():
    global _lib
    if _lib is None:
        if not LIB.exists():
            build()
        _lib = C.CDLL(str(LIB))
        _setup(_lib)
    return _lib


_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_up = np.ctypeslib.ndpointer(dtype=np.uint64, flags="C_CONTIGUOUS")


class Rng(C.Structure):
    """Xoshiro256++ state seeded by SplitMix64 (src/rng.rs:15-87)."""

    _fields_ = [("s", C.c_uint64 * 4)]

    def __init__(self, seed: int):
        super().__init__()
        lib().orc_rng_seed(C.byref(self), C.c_uint64(seed))

    def next_u64(self) -> int:
        return lib().orc_rng_next(C.byref(self))

    def unit(self, size: int | None = None):
        if size is None:
            return lib().orc_rng_unit(C.byref(self))
        out = np.empty(size)
        lib().orc_rng_fill_unit(C.byref(self), out, C.c_size_t(size))
        return out

    def normal(self, size: int | None = None):
        if size is None:
            return lib().orc_rng_normal(C.byref(self))
        out = np.empty(size)
        lib().orc_rng_fill_normal(C.byref(self), out, C.c_size_t(size))
        return out

    def up_to(self, mx: int) -> int:
        return lib().orc_rng_up_to(C.byref(self), C.c_uint64(mx))

    def shuffle(self, idx: np.ndarray) -> None:
        assert idx.dtype == np.uint64 and idx.flags.c_contiguous
        lib().orc_shuffle(C.byref(self), idx, C.c_size_t(idx.size))


def _setup(L):
    L.orc_rng_seed.argtypes = [C.c_void_p, C.c_uint64]
    L.orc_rng_next.argtypes = [C.c_void_p]
    L.orc_rng_next.restype = C.c_uint64
    L.orc_rng_unit.argtypes = [C.c_void_p]
    L.orc_rng_unit.restype = C.c_double
    L.orc_rng_normal.argtypes = [C.c_void_p]
    L.orc_rng_normal.restype = C.c_double
    L.orc_rng_up_to.argtypes = [C.c_void_p, C.c_uint64]
    L.orc_rng_up_to.restype = C.c_uint64
    L.orc_rng_fill_unit.argtypes = [C.c_void_p, _dp, C.c_size_t]
    L.orc_rng_fill_normal.argtypes = [C.c_void_p, _dp, C.c_size_t]
    L.orc_shuffle.argtypes = [C.c_void_p, _up, C.c_size_t]
    L.orc_num_params.argtypes = [_ip, C.c_int, C.c_int]
    L.orc_cross.argtypes = [_ip, C.c_int, _dp, _dp, C.c_size_t, _dp, C.c_size_t, C.c_int, _dp]
    L.orc_contract.argtypes = [_ip, C.c_int, _dp, _dp, C.c_size_t, _dp, C.c_size_t, C.c_int, _dp,
                               C.c_int, _dp]
    L.orc_exact_value_grad.argtypes = [_ip, C.c_int, _dp, _dp, C.c_size_t, C.c_int, _dp,
                                       C.c_double, C.c_double, C.c_int, C.c_double,
                                       C.POINTER(C.c_double), C.c_void_p, C.c_void_p,
                                       C.POINTER(C.c_double), C.c_int]
    L.orc_exact_predict.argtypes = [_ip, C.c_int, _dp, _dp, C.c_size_t, C.c_int, _dp, _dp,
                                    C.c_size_t, _dp, _dp, _dp]
    L.orc_vfe_value_grad.argtypes = [_ip, C.c_int, _dp, _dp, C.c_size_t, C.c_int, _dp, _dp,
                                     C.c_size_t, C.POINTER(C.c_double), C.c_void_p]
    L.orc_vfe_predict.argtypes = [_ip, C.c_int, _dp, _dp, C.c_size_t, C.c_int, _dp, _dp,
                                  C.c_size_t, _dp, C.c_size_t, _dp, _dp, _dp]
    L.orc_svgp_value_grad.argtypes = [_ip, C.c_int, _dp, _dp, C.c_size_t, C.c_int, _dp, _dp,
                                      C.c_size_t, _up, C.c_size_t, C.POINTER(C.c_double), _dp]
    L.orc_svgp_predict.argtypes = [_ip, C.c_int, _dp, C.c_int, _dp, C.c_size_t, _dp, C.c_size_t,
                                   _dp, _dp, _dp]
    L.orc_adam_step.argtypes = [_dp, _dp, _dp, _dp, C.POINTER(C.c_uint64), C.c_size_t,
                                C.c_double, C.c_double, C.c_double, C.c_double]
    L.orc_threads.restype = C.c_int
    L.orc_set_verbose.argtypes = [C.c_int]
    L.orc_set_verbose.restype = None


class OracleError(RuntimeError):
    def __init__(self, code: int, what: str):
        super().__init__(f"{what}: oracle status {code}")
        self.code = code


def _ops(ops) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(ops, dtype=np.int32).reshape(-1))


def _f(a) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a, dtype=np.float64).reshape(-1))


def num_params(ops, d: int) -> int:
    o = _ops(ops)
    return lib().orc_num_params(o, o.size // 2, d)


def cross(ops, theta, a, na, b, nb, d):
    o, out = _ops(ops), np.empty(na * nb)
    rc = lib().orc_cross(o, o.size // 2, _f(theta), _f(a), na, _f(b), nb, d, out)
    if rc:
        raise OracleError(rc, "cross")
    return out.reshape(nb, na).T


def contract(ops, theta, a, na, b, nb, d, V, sym=False):
    o = _ops(ops)
    p = num_params(o, d)
    out = np.empty(p)
    Vf = _f(np.asarray(V, dtype=np.float64).T)  # column-major
    rc = lib().orc_contract(o, o.size // 2, _f(theta), _f(a), na, _f(b), nb, d, Vf, int(sym), out)
    if rc:
        raise OracleError(rc, "contract")
    return out


def exact_value_grad(ops, theta, x, n, d, y, jitter=(0.0, 1.0, 0, 0.0), want_grad=True,
                     blocked=False):
    """NLML, gradient (or None), alpha and the jitter taken."""
    o = _ops(ops)
    p = num_params(o, d) + 1
    val = C.c_double()
    grad = np.empty(p)
    alpha = np.empty(n)
    taken = C.c_double()
    rc = lib().orc_exact_value_grad(o, o.size // 2, _f(theta), _f(x), n, d, _f(y),
                                    jitter[0], jitter[1], jitter[2], jitter[3], C.byref(val),
                                    grad.ctypes.data if want_grad else None, alpha.ctypes.data,
                                    C.byref(taken), int(blocked))
    if rc:
        raise OracleError(rc, "exact_value_grad")
    return val.value, (grad if want_grad else None), alpha, taken.value


def exact_predict(ops, theta, x, n, d, y, xs, q):
    o = _ops(ops)
    mean, lat, obs = np.empty(q), np.empty(q), np.empty(q)
    rc = lib().orc_exact_predict(o, o.size // 2, _f(theta), _f(x), n, d, _f(y), _f(xs), q, mean,
                                 lat, obs)
    if rc:
        raise OracleError(rc, "exact_predict")
    return mean, lat, obs


def vfe_value_grad(ops, theta, x, n, d, y, z, m, want_grad=True):
    o = _ops(ops)
    p = num_params(o, d) + 1
    val = C.c_double()
    grad = np.empty(p)
    rc = lib().orc_vfe_value_grad(o, o.size // 2, _f(theta), _f(x), n, d, _f(y), _f(z), m,
                                  C.byref(val), grad.ctypes.data if want_grad else None)
    if rc:
        raise OracleError(rc, "vfe_value_grad")
    return val.value, (grad if want_grad else None)


def vfe_predict(ops, theta, x, n, d, y, z, m, xs, q):
    o = _ops(ops)
    mean, lat, obs = np.empty(q), np.empty(q), np.empty(q)
    rc = lib().orc_vfe_predict(o, o.size // 2, _f(theta), _f(x), n, d, _f(y), _f(z), m, _f(xs), q,
                               mean, lat, obs)
    if rc:
        raise OracleError(rc, "vfe_predict")
    return mean, lat, obs


def svgp_value_grad(ops, params, x, n, d, y, z, m, batch):
    o = _ops(ops)
    b = np.ascontiguousarray(np.asarray(batch, dtype=np.uint64))
    val = C.c_double()
    grad = np.empty(len(params))
    rc = lib().orc_svgp_value_grad(o, o.size // 2, _f(params), _f(x), n, d, _f(y), _f(z), m, b,
                                   b.size, C.byref(val), grad)
    if rc:
        raise OracleError(rc, "svgp_value_grad")
    return val.value, grad


def svgp_predict(ops, params, d, z, m, xs, q):
    o = _ops(ops)
    mean, lat, obs = np.empty(q), np.empty(q), np.empty(q)
    rc = lib().orc_svgp_predict(o, o.size // 2, _f(params), d, _f(z), m, _f(xs), q, mean, lat, obs)
    if rc:
        raise OracleError(rc, "svgp_predict")
    return mean, lat, obs


def adam_step(z, g, m1, m2, t, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
    """In-place Kingma-Ba step; returns the new timestep."""
    tt = C.c_uint64(t)
    lib().orc_adam_step(z, _f(g), m1, m2, C.byref(tt), z.size, lr, b1, b2, eps)
    return tt.value


def threads() -> int:
    return lib().orc_threads()


def set_verbose(on: bool) -> None:
    """Print one line per finished phase of an exact evaluation to stderr."""
    lib().orc_set_verbose(int(on))
