"""Many independent exact-GPR problems evaluated together (one grouped launch
per factor step for all of them, replayed through a hipGraph).

    batch = GprBatch([(kernel, likelihood, x, n, d, y), ...])
    nlml, grads, errors = batch.value_and_gradient(thetas)

Each problem is what gprx's FittedGpr::value_and_gradient_into evaluates
(src/gpr/fitted.rs:620); the batch replaces looping over them, e.g. the
multi-start candidates of Lbfgs::with_restarts (src/optimizer/mod.rs:182-227).
"""
from __future__ import annotations

import ctypes as C
from typing import List, Sequence

import numpy as np

from . import _lib as L
from .errors import _BY_STATUS, GprError, raise_for
from .gpr import GaussianLikelihood, Precision
from .kernels import KernelSpec


class GprBatch:
    def __init__(self, problems: Sequence[tuple], precision: int = Precision.FP64):
        self._keep = []
        arr = (L.ProblemC * len(problems))()
        self.p = []
        for i, (kernel, lik, x, n, d, y) in enumerate(problems):
            x, y = L.f64(x), L.f64(y)
            ops = np.ascontiguousarray(kernel.program(), dtype=np.int32)
            prog = L.KernelProgram(ops.ctypes.data_as(C.POINTER(C.c_int32)), ops.size // 2)
            th = np.concatenate([kernel.get_params(), [lik.log_param()]])
            self._keep += [x, y, ops, prog, th]
            arr[i] = L.ProblemC(L.dptr(x), n, d, L.dptr(y), C.pointer(prog), L.dptr(th), th.size)
            self.p.append(th.size)
        err = L.GprxError()
        h = C.c_void_p()
        rc = L.lib().gprx_hip_batch_create(arr, len(problems), precision, C.byref(h), C.byref(err))
        raise_for(rc, err)
        self._h = h
        self.k = len(problems)

    def __del__(self):
        self.close()

    def close(self):
        if getattr(self, "_h", None):
            L.lib().gprx_hip_destroy(self._h)
            self._h = None

    def value_and_gradient(self, thetas: Sequence[np.ndarray]):
        """(nlml[k], [grad_i], [GprError or None])."""
        ths = [L.f64(t) for t in thetas]
        tptr = (L._dp * self.k)(*[L.dptr(t) for t in ths])
        grads = [np.empty(p) for p in self.p]
        gptr = (L._dp * self.k)(*[L.dptr(g) for g in grads])
        nlml = np.empty(self.k)
        errs = (L.GprxError * self.k)()
        rc = L.lib().gprx_hip_batch_value_and_grad(self._h, tptr, L.dptr(nlml), gptr, errs)
        raise_for(rc, errs[0])
        out_err: List = []
        for e in errs:
            if e.status:
                out_err.append(_BY_STATUS.get(e.status, GprError)(e.message.decode(errors="replace")))
            else:
                out_err.append(None)
        return nlml, grads, out_err

    def set_schedule(self, policy: int, small_tile_below: int = 0) -> None:
        """Fixes the launch schedule (policy 0, 1 or 2; -1 returns to the
        autotuner), see gprx_hip_batch_set_schedule."""
        err = L.GprxError()
        rc = L.lib().gprx_hip_batch_set_schedule(self._h, policy, small_tile_below, C.byref(err))
        raise_for(rc, err)

    def schedule(self):
        """(policy, small_tile_below, tuned_ms) in use; policy -1 before the
        first evaluation."""
        pol, below, ms = C.c_int32(), C.c_int32(), C.c_double()
        L.lib().gprx_hip_batch_schedule(self._h, C.byref(pol), C.byref(below), C.byref(ms))
        return pol.value, below.value, ms.value

    def last_fallback(self) -> np.ndarray:
        """Per problem, why the last mixed call evaluated it alone in FP64
        (FALLBACK_NONE 0, FALLBACK_REFINEMENT 1, FALLBACK_JITTER 2)."""
        out = np.zeros(self.k, dtype=np.int32)
        rc = L.lib().gprx_hip_batch_last_fallback(self._h, out.ctypes.data_as(C.POINTER(C.c_int32)),
                                                  self.k)
        raise_for(rc, L.GprxError())
        return out

    def last_eval_ms(self) -> float:
        return L.lib().gprx_hip_last_eval_ms(self._h)
