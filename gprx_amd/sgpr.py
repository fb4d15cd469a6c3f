"""Sparse VFE GPR (Sgpr) on the gfx950 backend, mirroring gprx's Sgpr API.

    fitted = Sgpr(kernel, GaussianLikelihood(0.1)).factor(x, n, d, y, z, m)
    value = fitted.value_and_gradient_into(params, grad)

Reference interfaces: Sgpr::factor (src/sgpr/model.rs:448-513),
FittedSgpr::value_and_gradient_into (src/sgpr/fitted.rs:306-326),
FittedSgpr::predict_with (src/sgpr/factor/predict.rs:102-260).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as L
from .errors import raise_for
from .gpr import GaussianLikelihood, Precision, Prediction, VarianceKind, program_c
from .kernels import KernelSpec


class Sgpr:
    def __init__(self, kernel: KernelSpec, likelihood: GaussianLikelihood):
        self.kernel = kernel
        self.likelihood = likelihood
        self.precision = Precision.FP64

    def with_precision(self, precision: int) -> "Sgpr":
        self.precision = precision
        return self

    def prepare(self, x, n: int, d: int, y, z, m: int) -> "FittedSgpr":
        x, y, z = L.f64(x), L.f64(y), L.f64(z)
        if x.size != n * d or y.size != n or z.size != m * d:
            raise_for(L.LENGTH_MISMATCH, "x, y, z sizes differ from n, d, m")
        theta = np.concatenate([self.kernel.get_params(), [self.likelihood.log_param()]])
        prog, ops = program_c(self.kernel)
        err = L.GprxError()
        h = C.c_void_p()
        rc = L.lib().gprx_hip_sgpr_create(L.dptr(x), n, d, L.dptr(y), L.dptr(z), m, C.byref(prog),
                                          L.dptr(theta), theta.size, self.precision, C.byref(h),
                                          C.byref(err))
        raise_for(rc, err)
        f = FittedSgpr(h, theta, n, d, m)
        # what a save writes back (x and z column-major, d rows of n and m)
        f._kernel, f._precision = self.kernel, self.precision
        f._x, f._y, f._z = x.copy(), y.copy(), z.copy()
        return f

    def factor(self, x, n, d, y, z, m) -> "FittedSgpr":
        f = self.prepare(x, n, d, y, z, m)
        f.neg_log_marginal_likelihood()
        return f


class FittedSgpr:
    def __init__(self, h, theta, n, d, m):
        self._h, self._theta = h, theta.copy()
        self._p = theta.size
        self.n, self.d, self.m = n, d, m

    def __del__(self):
        self.close()

    def close(self):
        if getattr(self, "_h", None):
            L.lib().gprx_hip_destroy(self._h)
            self._h = None

    def num_params(self) -> int:
        return self._p

    def get_params(self) -> np.ndarray:
        return self._theta.copy()

    def value_and_gradient_into(self, params, out) -> float:
        p = L.f64(params)
        g = np.empty(self._p)
        val = C.c_double()
        err = L.GprxError()
        rc = L.lib().gprx_hip_sgpr_value_and_grad(self._h, L.dptr(p), p.size, C.byref(val),
                                                  L.dptr(g), C.byref(err))
        raise_for(rc, err)
        self._theta = p.copy()
        out[:] = g
        return val.value

    def value(self, params) -> float:
        p = L.f64(params)
        val = C.c_double()
        err = L.GprxError()
        rc = L.lib().gprx_hip_sgpr_value_and_grad(self._h, L.dptr(p), p.size, C.byref(val), None,
                                                  C.byref(err))
        raise_for(rc, err)
        self._theta = p.copy()
        return val.value

    def neg_log_marginal_likelihood(self) -> float:
        return self.value(self._theta)

    def predict(self, xs, q, d) -> Prediction:
        return self.predict_with(xs, q, d, VarianceKind.OBSERVATION)

    def predict_with(self, xs, q, d, variance_kind=VarianceKind.OBSERVATION) -> Prediction:
        xs = L.f64(xs)
        mean, var = np.empty(q), np.empty(q)
        err = L.GprxError()
        rc = L.lib().gprx_hip_sgpr_predict(self._h, L.dptr(xs), q, d, variance_kind, L.dptr(mean),
                                           L.dptr(var), C.byref(err))
        raise_for(rc, err)
        return Prediction(mean, var)

    def save(self, directory) -> None:
        """save_sgpr (src/persist/sparse.rs:165-175): config.json and the x,
        y, z tensors at the current parameters; no factor is written."""
        from .persist import save_sparse
        save_sparse("sgpr", self._kernel, self._theta, self._x, self._y, self._z, self.n,
                    self.m, self.d, self._precision == L.MIXED, directory)

    @staticmethod
    def load(directory) -> "FittedSgpr":
        """LoadedSgpr::load (src/persist/sparse.rs:547-556) for a coordinate
        Sgpr: reads and checks the directory on the host, then factors on the
        device at the saved parameters and inducing points."""
        from .persist import read_sparse
        kernel, noise, mixed, x, y, z, n, m, d, _ = read_sparse(directory, "sgpr")
        model = Sgpr(kernel, GaussianLikelihood(noise))
        if mixed:
            model = model.with_precision(L.MIXED)
        return model.factor(x, n, d, y, z, m)

    def last_eval_ms(self) -> float:
        return L.lib().gprx_hip_last_eval_ms(self._h)
