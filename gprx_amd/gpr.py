"""Exact GPR on the gfx950 backend, mirroring gprx's Gpr / FittedGpr API.

    fitted = Gpr(kernel, GaussianLikelihood(0.1)).factor(x, n, d, y)
    value = fitted.value_and_gradient_into(params, grad)      # objective seam
    pred = fitted.predict(xs, q, d)                            # mean, variance

Reference interfaces: Gpr::factor (src/gpr/trainer.rs:533-630),
FittedGpr::value_and_gradient_into (src/gpr/fitted.rs:620),
FittedGpr::predict / predict_with (src/gpr/fitted.rs:854-960).
"""
from __future__ import annotations

import copy
import ctypes as C
import math
from dataclasses import dataclass

import numpy as np

from . import _lib as L
from .errors import InvalidConfig, raise_for
from .kernels import KernelSpec


class GaussianLikelihood:
    """Noise variance sigma^2; its parameter is log sigma^2 (src/likelihood.rs:57-104)."""

    def __init__(self, noise_variance: float):
        if not (noise_variance > 0 and math.isfinite(noise_variance)):
            raise ValueError("noise variance must be positive and finite")
        self.noise_variance = float(noise_variance)

    def log_param(self) -> float:
        return math.log(self.noise_variance)


@dataclass(frozen=True)
class JitterPolicy:
    """gprx's JitterPolicy (src/policy.rs:180-396)."""

    initial: float = 0.0
    multiplier: float = 1.0
    max_jitter: float = 0.0
    max_retries: int = 0

    @staticmethod
    def fixed(j: float) -> "JitterPolicy":
        """JitterPolicy::fixed (src/policy.rs:245-252): j finite and >= 0."""
        if not (math.isfinite(j) and j >= 0.0):
            raise InvalidConfig(f"jitter must be finite and non-negative, got {j}")
        return JitterPolicy(j, 1.0, j, 1 if j > 0 else 0)

    @staticmethod
    def adaptive(initial: float, multiplier: float, max_retries: int, max_jitter: float):
        """JitterPolicy::adaptive (src/policy.rs:266-304): initial > 0,
        multiplier > 1, max_retries >= 1 and max_jitter >= initial, all finite."""
        if not (math.isfinite(initial) and initial > 0.0):
            raise InvalidConfig(f"adaptive initial jitter must be finite and positive, got {initial}")
        if not (math.isfinite(multiplier) and multiplier > 1.0):
            raise InvalidConfig("adaptive jitter multiplier must be finite and greater than 1, "
                                f"got {multiplier}")
        if max_retries < 1:
            raise InvalidConfig("adaptive jitter max_retries must be at least 1")
        if not (math.isfinite(max_jitter) and max_jitter >= initial):
            raise InvalidConfig(f"adaptive max_jitter must be finite and at least initial ({initial}), "
                                f"got {max_jitter}")
        return JitterPolicy(initial, multiplier, max_jitter, max_retries)

    def c(self) -> L.JitterPolicyC:
        return L.JitterPolicyC(self.initial, self.multiplier, self.max_jitter, self.max_retries)


class Precision:
    FP64 = L.FP64
    MIXED = L.MIXED


class Fallback:
    NONE = L.FALLBACK_NONE
    REFINEMENT = L.FALLBACK_REFINEMENT
    JITTER = L.FALLBACK_JITTER


class VarianceKind:
    OBSERVATION = L.VARIANCE_OBSERVATION
    LATENT = L.VARIANCE_LATENT


@dataclass
class Prediction:
    mean: np.ndarray
    variance: np.ndarray | None


def program_c(kernel: KernelSpec):
    ops = np.ascontiguousarray(kernel.program(), dtype=np.int32)
    prog = L.KernelProgram(ops.ctypes.data_as(C.POINTER(C.c_int32)), ops.size // 2)
    return prog, ops  # keep ops alive with prog


class Gpr:
    def __init__(self, kernel: KernelSpec, likelihood: GaussianLikelihood):
        self.kernel = kernel
        self.likelihood = likelihood
        self.jitter = JitterPolicy()
        self.precision = Precision.FP64
        self.mixed_jitter_bound = 0.0
        self.optimizer = None
        self.bounds = None

    def with_optimizer(self, optimizer) -> "Gpr":
        """Lbfgs or TrustRegion (with or without restarts), or Fixed, used by fit()."""
        self.optimizer = optimizer
        return self

    def with_bounds(self, bounds) -> "Gpr":
        """User-unit (lo, hi) of every parameter in get_params order (kernel
        leaves, then the noise variance); default (1e-5, 1e5) each."""
        self.bounds = None if bounds is None else L.f64(bounds)
        return self

    def fit(self, x, n: int, d: int, y) -> "FittedGpr":
        """Gpr::fit (src/gpr/trainer.rs): minimizes the NLML with the
        optimizer (default Lbfgs()), then factors at the best parameters.
        The fitted model's fit_result holds the optimizer's OptResult."""
        from .optimizer import Fixed, Lbfgs, OptResult, TrustRegion
        opt = self.optimizer if self.optimizer is not None else Lbfgs()
        if isinstance(opt, Fixed):
            return self.factor(x, n, d, y)
        if isinstance(opt, TrustRegion):
            fitted = self.prepare(x, n, d, y)
            th0 = fitted.get_params()
            if self.bounds is not None and self.bounds.size != 2 * th0.size:
                raise_for(L.LENGTH_MISMATCH, "bounds must hold (lo, hi) for every parameter")
            best = np.empty(th0.size)
            val = C.c_double()
            it = C.c_uint64()
            start = C.c_int32()
            err = L.GprxError()
            oc = opt.options_c()
            bptr = L.dptr(self.bounds) if self.bounds is not None else None
            rc = L.lib().gprx_hip_exact_fit_trust_region(fitted._h, L.dptr(th0), th0.size, bptr,
                                                         C.byref(oc), L.dptr(best), C.byref(val),
                                                         C.byref(it), C.byref(start),
                                                         C.byref(err))
            raise_for(rc, err)
            fitted.fit_result = OptResult(best.copy(), val.value, it.value, start.value)
            return fitted
        x = L.f64(x)
        y = L.f64(y)
        if x.size != n * d or y.size != n:
            raise_for(L.LENGTH_MISMATCH, "x must hold n*d values and y n values")
        theta = np.concatenate([self.kernel.get_params(), [self.likelihood.log_param()]])
        if self.bounds is not None and self.bounds.size != 2 * theta.size:
            raise_for(L.LENGTH_MISMATCH, "bounds must hold (lo, hi) for every parameter")
        prog, ops = program_c(self.kernel)
        best = np.empty(theta.size)
        val = C.c_double()
        it = C.c_uint64()
        start = C.c_int32()
        err = L.GprxError()
        oc = opt.options_c()
        bptr = L.dptr(self.bounds) if self.bounds is not None else None
        jit = self.jitter.c()
        rc = L.lib().gprx_hip_exact_fit(L.dptr(x), n, d, L.dptr(y), C.byref(prog), L.dptr(theta),
                                        theta.size, bptr, self.precision, C.byref(jit), C.byref(oc),
                                        L.dptr(best), C.byref(val), C.byref(it), C.byref(start),
                                        C.byref(err))
        raise_for(rc, err)
        fitted = self._create(x, n, d, y, best)
        fitted.fit_result = OptResult(best.copy(), val.value, it.value, start.value)
        fitted.neg_log_marginal_likelihood()
        return fitted

    def with_jitter_policy(self, policy: JitterPolicy) -> "Gpr":
        self.jitter = policy
        return self

    def with_precision(self, precision: int) -> "Gpr":
        self.precision = precision
        return self

    def with_mixed_jitter_bound(self, bound: float) -> "Gpr":
        """Largest jitter the mixed form's f32 factor may add before the
        evaluation is formed in FP64 instead (default 0)."""
        self.mixed_jitter_bound = bound
        return self

    def factor(self, x, n: int, d: int, y) -> "FittedGpr":
        """Gpr::<Fixed>::factor: the model at the spec's parameters, factored."""
        fitted = self.prepare(x, n, d, y)
        fitted.neg_log_marginal_likelihood()
        return fitted

    def prepare(self, x, n: int, d: int, y) -> "FittedGpr":
        """Uploads the data without forming a factor."""
        x = L.f64(x)
        y = L.f64(y)
        if x.size != n * d or y.size != n:
            raise_for(L.LENGTH_MISMATCH, "x must hold n*d values and y n values")
        theta = np.concatenate([self.kernel.get_params(), [self.likelihood.log_param()]])
        return self._create(x, n, d, y, theta)

    def _create(self, x, n: int, d: int, y, theta: np.ndarray) -> "FittedGpr":
        prog, ops = program_c(self.kernel)
        err = L.GprxError()
        handle = C.c_void_p()
        jit = self.jitter.c()
        rc = L.lib().gprx_hip_exact_create(L.dptr(x), n, d, L.dptr(y), C.byref(prog),
                                           L.dptr(theta), theta.size, self.precision,
                                           C.byref(jit), C.byref(handle), C.byref(err))
        raise_for(rc, err)
        fitted = FittedGpr(handle, theta, n, d)
        # what a save writes: this configuration and the training data
        # (x column-major, kept as d rows of n values)
        fitted._gpr = copy.copy(self)
        fitted._x = np.array(x, dtype=np.float64).reshape(d, n)
        fitted._y = np.array(y, dtype=np.float64).reshape(n)
        if self.mixed_jitter_bound != 0.0:
            rc = L.lib().gprx_hip_set_mixed_jitter_bound(handle, self.mixed_jitter_bound,
                                                         C.byref(err))
            raise_for(rc, err)
        return fitted


class FittedGpr:
    def __init__(self, handle: C.c_void_p, theta: np.ndarray, n: int, d: int):
        self._h = handle
        self._p = theta.size
        self.n, self.d = n, d
        self._theta = theta.copy()

    def __del__(self):
        self.close()

    def close(self):
        if getattr(self, "_h", None):
            L.lib().gprx_hip_destroy(self._h)
            self._h = None

    def num_params(self) -> int:
        return self._p

    def get_params(self, out=None) -> np.ndarray:
        buf = np.empty(self._p)
        raise_for(L.lib().gprx_hip_get_params(self._h, L.dptr(buf), self._p), None)
        if out is not None:
            out[:] = buf
        return buf

    def value_and_gradient_into(self, params, out) -> float:
        """NLML at log-params; writes the gradient into `out` (src/objective.rs:308-340)."""
        p = L.f64(params)
        g = np.empty(self._p)
        val = C.c_double()
        err = L.GprxError()
        rc = L.lib().gprx_hip_exact_value_and_grad(self._h, L.dptr(p), p.size, C.byref(val),
                                                   L.dptr(g), C.byref(err))
        raise_for(rc, err)
        out[:] = g
        return val.value

    def value_gradient_hessian_into(self, params, grad, hess) -> float:
        """TwiceDifferentiable::value_gradient_hessian_into: NLML, gradient and
        the p x p Hessian (FP64) at params; the model moves to params."""
        th = L.f64(params)
        if grad.size != th.size or hess.size != th.size * th.size:
            raise_for(L.LENGTH_MISMATCH, "grad must hold p values and hess p*p")
        g = np.empty(th.size)
        h = np.empty(th.size * th.size)
        v = C.c_double()
        err = L.GprxError()
        rc = L.lib().gprx_hip_exact_hessian(self._h, L.dptr(th), th.size, C.byref(v), L.dptr(g),
                                            L.dptr(h), C.byref(err))
        raise_for(rc, err)
        grad.reshape(-1)[:] = g
        hess.reshape(-1)[:] = h
        return v.value

    def insert(self, x_new, y_new: float) -> float:
        """OnlineGpr::insert: appends one point and its target, updating the
        FP64 factor in O(n^2); returns the new NLML."""
        xn = L.f64(x_new)
        v = C.c_double()
        err = L.GprxError()
        rc = L.lib().gprx_hip_exact_insert(self._h, L.dptr(xn), xn.size, float(y_new), C.byref(v),
                                           C.byref(err))
        raise_for(rc, err)
        self.n += 1
        self._x = np.hstack([self._x, xn.reshape(self.d, 1)])
        self._y = np.append(self._y, float(y_new))
        return v.value

    def delete(self, index: int) -> float:
        """OnlineGpr::delete: drops the point at row `index` (later rows move
        up by one) and factors again; returns the new NLML."""
        v = C.c_double()
        err = L.GprxError()
        rc = L.lib().gprx_hip_exact_delete(self._h, int(index), C.byref(v), C.byref(err))
        raise_for(rc, err)
        self.n -= 1
        self._x = np.delete(self._x, int(index), axis=1)
        self._y = np.delete(self._y, int(index))
        return v.value

    def save(self, directory) -> None:
        """FittedGpr::save (src/gpr/fitted.rs:810-812): config.json and the
        training x, y in model.safetensors, written atomically."""
        from .persist import save_fitted
        save_fitted(self, directory, False)

    def save_with_factor(self, directory) -> None:
        """FittedGpr::save_with_factor (src/gpr/fitted.rs:814-830): as save,
        with the Cholesky factor l (column-major) and alpha, formed on the
        device in FP64 (f32 for a mixed model)."""
        from .persist import save_fitted
        save_fitted(self, directory, True)

    @staticmethod
    def load(directory) -> "FittedGpr":
        """LoadedGpr::load (src/persist/mod.rs:308-310) of an exact model."""
        from .persist import load
        return load(directory)

    def value(self, params) -> float:
        p = L.f64(params)
        val = C.c_double()
        err = L.GprxError()
        rc = L.lib().gprx_hip_exact_value_and_grad(self._h, L.dptr(p), p.size, C.byref(val),
                                                   None, C.byref(err))
        raise_for(rc, err)
        return val.value

    def neg_log_marginal_likelihood(self) -> float:
        return self.value(self.get_params())

    def predict(self, xs, q: int, d: int) -> Prediction:
        return self.predict_with(xs, q, d, VarianceKind.OBSERVATION)

    def predict_with(self, xs, q: int, d: int, variance_kind: int = VarianceKind.OBSERVATION,
                     mean_only: bool = False) -> Prediction:
        xs = L.f64(xs)
        if xs.size != q * d:
            raise_for(L.LENGTH_MISMATCH, "xs must hold q*d values")
        mean = np.empty(q)
        var = None if mean_only else np.empty(q)
        err = L.GprxError()
        rc = L.lib().gprx_hip_exact_predict(self._h, L.dptr(xs), q, d, variance_kind,
                                            L.dptr(mean), None if var is None else L.dptr(var),
                                            C.byref(err))
        raise_for(rc, err)
        return Prediction(mean, var)

    # Diagnostics of the last evaluation.
    def last_jitter(self) -> float:
        return L.lib().gprx_hip_last_jitter(self._h)

    def last_fallback(self) -> int:
        """Why the last mixed evaluation ran in FP64: FALLBACK_NONE (0),
        FALLBACK_REFINEMENT (1) or FALLBACK_JITTER (2)."""
        return L.lib().gprx_hip_last_fallback(self._h)

    def last_refine_steps(self) -> int:
        return L.lib().gprx_hip_last_refine_steps(self._h)

    def last_eval_ms(self) -> float:
        return L.lib().gprx_hip_last_eval_ms(self._h)

    def last_dense_ms(self) -> float:
        return L.lib().gprx_hip_last_dense_ms(self._h)

    def last_phase_ms(self) -> list:
        """[fill, factor+inverse, alpha, X^T X, contraction] device ms."""
        return [L.lib().gprx_hip_last_phase_ms(self._h, k) for k in range(5)]
