"""Minibatch SVGP on the gfx950 backend, mirroring gprx's Svgp API.

    model = Svgp(kernel, GaussianLikelihood(0.1)).prepare(x, n, d, y, z, m)
    value = model.value_and_gradient_into(params, grad, batch)
    params, values = model.adam(params, steps=2000, batch_size=4096, lr=0.01)

Parameters are [log theta | log s2 | mu (m) | packed Lq (columns, i >= j)]
(src/svgp/factor/assemble.rs:158-225).  Reference interfaces:
svgp_value_and_gradient_with (src/svgp/factor/gradient.rs:104-249),
run_adam_fit (src/svgp/factor/adam.rs:92-141).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as L
from .errors import raise_for
from .gpr import GaussianLikelihood, Precision, Prediction, VarianceKind, program_c
from .kernels import KernelSpec


def prior_params(theta: np.ndarray, m: int) -> np.ndarray:
    """theta followed by the prior q(u): mu = 0, Lq = I (src/svgp/factor/assemble.rs:144-156)."""
    packed = np.zeros(m * (m + 1) // 2)
    k = 0
    for j in range(m):
        packed[k] = 1.0
        k += m - j
    return np.concatenate([theta, np.zeros(m), packed])


class Svgp:
    def __init__(self, kernel: KernelSpec, likelihood: GaussianLikelihood):
        self.kernel = kernel
        self.likelihood = likelihood
        self.precision = Precision.FP64

    def with_precision(self, precision: int) -> "Svgp":
        self.precision = precision
        return self

    def prepare(self, x, n, d, y, z, m) -> "FittedSvgp":
        x, y, z = L.f64(x), L.f64(y), L.f64(z)
        if x.size != n * d or y.size != n or z.size != m * d:
            raise_for(L.LENGTH_MISMATCH, "x, y, z sizes differ from n, d, m")
        theta = np.concatenate([self.kernel.get_params(), [self.likelihood.log_param()]])
        prog, ops = program_c(self.kernel)
        err = L.GprxError()
        h = C.c_void_p()
        rc = L.lib().gprx_hip_svgp_create(L.dptr(x), n, d, L.dptr(y), L.dptr(z), m, C.byref(prog),
                                          L.dptr(theta), theta.size, self.precision, C.byref(h),
                                          C.byref(err))
        raise_for(rc, err)
        f = FittedSvgp(h, theta, n, d, m)
        # what a save writes back (x and z column-major, d rows of n and m)
        f._kernel, f._precision = self.kernel, self.precision
        f._x, f._y, f._z = x.copy(), y.copy(), z.copy()
        return f


class FittedSvgp:
    def __init__(self, h, theta, n, d, m):
        self._h = h
        self.theta = theta.copy()
        self.n, self.d, self.m = n, d, m
        self.params = prior_params(theta, m)

    def __del__(self):
        self.close()

    def close(self):
        if getattr(self, "_h", None):
            L.lib().gprx_hip_destroy(self._h)
            self._h = None

    def num_params(self) -> int:
        return int(L.lib().gprx_hip_svgp_num_params(self._h))

    def value_and_gradient_into(self, params, out, batch=None) -> float:
        p = L.f64(params)
        b = np.arange(self.n, dtype=np.uint64) if batch is None else np.ascontiguousarray(batch, dtype=np.uint64)
        g = np.empty(p.size)
        val = C.c_double()
        err = L.GprxError()
        rc = L.lib().gprx_hip_svgp_value_and_grad(self._h, L.dptr(p), p.size,
                                                  b.ctypes.data_as(C.POINTER(C.c_uint64)), b.size,
                                                  C.byref(val), L.dptr(g), C.byref(err))
        raise_for(rc, err)
        out[:] = g
        return val.value

    def adam(self, params=None, steps=100, batch_size=32, lr=1e-3, beta1=0.9, beta2=0.999,
             eps=1e-8, seed=0):
        """`steps` Adam steps (src/optimizer/adam.rs:179-198) over the reference's
        epoch shuffle; returns (params, -ELBO per step)."""
        p = L.f64(self.params if params is None else params).copy()
        vals = np.empty(steps)
        err = L.GprxError()
        rc = L.lib().gprx_hip_svgp_adam(self._h, L.dptr(p), p.size, batch_size, steps, lr, beta1,
                                        beta2, eps, seed, L.dptr(vals), C.byref(err))
        raise_for(rc, err)
        self.params = p
        return p, vals

    def predict_with(self, xs, q, d, params=None, variance_kind=VarianceKind.OBSERVATION) -> Prediction:
        p = L.f64(self.params if params is None else params)
        xs = L.f64(xs)
        mean, var = np.empty(q), np.empty(q)
        err = L.GprxError()
        rc = L.lib().gprx_hip_svgp_predict(self._h, L.dptr(p), p.size, L.dptr(xs), q, d,
                                           variance_kind, L.dptr(mean), L.dptr(var), C.byref(err))
        raise_for(rc, err)
        return Prediction(mean, var)

    def predict(self, xs, q, d, params=None) -> Prediction:
        return self.predict_with(xs, q, d, params)

    def save(self, directory, params=None) -> None:
        """save_svgp (src/persist/sparse.rs:190-195): config.json and the x,
        y, z, q_mean and q_l tensors of `params` (the model's own by default);
        no factor is written."""
        from .persist import save_sparse
        p = L.f64(self.params if params is None else params)
        t, m = self.theta.size, self.m
        if p.size != t + m + m * (m + 1) // 2:
            raise_for(L.LENGTH_MISMATCH, f"params has {p.size} entries, the model takes "
                                         f"{t + m + m * (m + 1) // 2}")
        save_sparse("svgp", self._kernel, p[:t], self._x, self._y, self._z, self.n, m, self.d,
                    self._precision == L.MIXED, directory, q=(p[t:t + m], p[t + m:]))

    @staticmethod
    def load(directory) -> "FittedSvgp":
        """LoadedSvgp::load (src/persist/sparse.rs:728-737) for a coordinate
        Svgp: reads and checks the directory on the host, then prepares the
        device model with the saved parameters and q(u) as its own."""
        from .persist import read_sparse
        kernel, noise, mixed, x, y, z, n, m, d, q = read_sparse(directory, "svgp")
        model = Svgp(kernel, GaussianLikelihood(noise))
        if mixed:
            model = model.with_precision(L.MIXED)
        f = model.prepare(x, n, d, y, z, m)
        f.params = np.concatenate([f.theta, q[0], q[1]])
        return f

    def last_eval_ms(self) -> float:
        return L.lib().gprx_hip_last_eval_ms(self._h)
