"""Optimizers of the exact path's hyperparameters, mirroring gprx's Lbfgs / Fixed.

    fitted = Gpr(kernel, lik).with_optimizer(Lbfgs().with_restarts(3, seed=0)).fit(x, n, d, y)

Reference interfaces: Lbfgs and its builder (src/optimizer/lbfgs.rs:19-116),
TrustRegion (src/optimizer/trust_region.rs:19-148),
the restart driver (src/optimizer/mod.rs:182-227), Fixed (mod.rs:150-170),
OptResult (mod.rs:40-60).  The solvers run in libgprx_hip.so
(gprx_lbfgs_minimize / gprx_hip_exact_fit; gprx_trust_region_minimize /
gprx_hip_exact_fit_trust_region); every start of an L-BFGS fit is
evaluated together with the others by the batched GPU path, and every
trust-region point takes its Hessian from the GPU.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass

import numpy as np

from . import _lib as L
from .errors import InvalidConfig, raise_for


@dataclass
class OptResult:
    params: np.ndarray
    value: float
    iterations: int
    best_start: int = 0


class Fixed:
    """Keeps the spec's parameters (factor only)."""


class Lbfgs:
    """L-BFGS with a More-Thuente line search in logit coordinates."""

    def __init__(self):
        self.max_iterations = 100
        self.tolerance = math.sqrt(np.finfo(float).eps)
        self.history_size = 10
        self.restarts = 0
        self.seed = 0

    def with_max_iterations(self, n: int) -> "Lbfgs":
        self.max_iterations = int(n)
        return self

    def with_tolerance(self, tolerance: float) -> "Lbfgs":
        if not math.isfinite(tolerance) or tolerance < 0.0:
            raise InvalidConfig("L-BFGS gradient tolerance must be finite and >= 0")
        self.tolerance = float(tolerance)
        return self

    def with_history_size(self, m: int) -> "Lbfgs":
        if m < 1:
            raise InvalidConfig("L-BFGS history size must be at least 1")
        self.history_size = int(m)
        return self

    def with_restarts(self, n: int, seed: int) -> "Lbfgs":
        if n < 1:
            raise InvalidConfig("restarts must be at least 1")
        self.restarts = int(n)
        self.seed = int(seed)
        return self

    def options_c(self) -> L.LbfgsOptionsC:
        return L.LbfgsOptionsC(self.max_iterations, self.tolerance, self.history_size,
                               self.restarts, self.seed)

    def minimize_unconstrained(self, fn, x0) -> OptResult:
        """Minimizes fn(x) -> (value, gradient) from x0 with this solver (host only;
        no restarts, no logit map)."""
        x0 = L.f64(x0)
        n = x0.size

        def cb(_ctx, xp, nn, vp, gp):
            try:
                x = np.ctypeslib.as_array(xp, shape=(nn,)).copy()
                v, g = fn(x)
                vp[0] = float(v)
                g = np.asarray(g, dtype=np.float64).reshape(-1)
                for i in range(nn):
                    gp[i] = g[i]
                return 0
            except Exception:  # a failed point: the solver treats it as a barrier
                return 1

        cfn = L.ObjectiveFn(cb)
        out = np.empty(n)
        val = C.c_double()
        it = C.c_uint64()
        err = L.GprxError()
        opt = self.options_c()
        rc = L.lib().gprx_lbfgs_minimize(cfn, None, L.dptr(x0), n, C.byref(opt), L.dptr(out),
                                         C.byref(val), C.byref(it), C.byref(err))
        raise_for(rc, err)
        return OptResult(out, val.value, it.value)


class TrustRegion:
    """Trust region with a Steihaug subproblem and the analytic Hessian."""

    def __init__(self):
        self.max_iterations = 100
        self.tolerance = math.sqrt(np.finfo(float).eps)
        self.initial_radius = 1.0
        self.max_radius = 100.0
        self.restarts = 0
        self.seed = 0

    def with_max_iterations(self, n: int) -> "TrustRegion":
        self.max_iterations = int(n)
        return self

    def with_tolerance(self, tolerance: float) -> "TrustRegion":
        if not math.isfinite(tolerance) or tolerance < 0.0:
            raise InvalidConfig("trust-region gradient tolerance must be finite and >= 0")
        self.tolerance = float(tolerance)
        return self

    def with_radii(self, initial: float, largest: float) -> "TrustRegion":
        if not (math.isfinite(initial) and math.isfinite(largest) and 0.0 < initial <= largest):
            raise InvalidConfig("trust-region radii need 0 < initial <= largest")
        self.initial_radius, self.max_radius = float(initial), float(largest)
        return self

    def with_restarts(self, n: int, seed: int) -> "TrustRegion":
        if n < 1:
            raise InvalidConfig("restarts must be at least 1")
        self.restarts = int(n)
        self.seed = int(seed)
        return self

    def options_c(self) -> L.TrOptionsC:
        return L.TrOptionsC(self.max_iterations, self.tolerance, self.initial_radius,
                            self.max_radius, self.restarts, self.seed)

    def minimize_unconstrained(self, fn, x0) -> OptResult:
        """Minimizes fn(x) -> (value, gradient, Hessian) from x0 (host only)."""
        x0 = L.f64(x0)
        n = x0.size

        def cb(_ctx, xp, nn, vp, gp, hp):
            try:
                x = np.ctypeslib.as_array(xp, shape=(nn,)).copy()
                v, g, h = fn(x)
                vp[0] = float(v)
                g = np.asarray(g, dtype=np.float64).reshape(-1)
                h = np.asarray(h, dtype=np.float64).reshape(-1)
                for i in range(nn):
                    gp[i] = g[i]
                for i in range(nn * nn):
                    hp[i] = h[i]
                return 0
            except Exception:  # a failed point
                return 1

        cfn = L.Objective2Fn(cb)
        out = np.empty(n)
        val = C.c_double()
        it = C.c_uint64()
        err = L.GprxError()
        opt = self.options_c()
        rc = L.lib().gprx_trust_region_minimize(cfn, None, L.dptr(x0), n, C.byref(opt),
                                                L.dptr(out), C.byref(val), C.byref(it),
                                                C.byref(err))
        raise_for(rc, err)
        return OptResult(out, val.value, it.value)
