"""KernelSpec trees, mirroring gprx's kernel API (src/kernel/spec.rs:106-132).

A spec is a tree of leaves combined with `+` and `*`.  `program()` flattens
it into the postfix op list of the C ABI (include/gprx_hip.h): associative
chains of the same operator collapse into one n-ary node, as gprx's
`compile()` does (src/kernel/spec.rs:384-416).  `get_params()` returns the
log-parameters in depth-first leaf order (src/kernel/spec.rs:29).
"""
from __future__ import annotations

import math
from typing import List, Sequence

import numpy as np

OP_CONST, OP_RBF, OP_RBF_ARD, OP_PERIODIC, OP_SUM, OP_PROD = 1, 2, 3, 4, 5, 6


class KernelSpec:
    def __add__(self, other: "KernelSpec") -> "KernelSpec":
        return Sum(_children(self, Sum) + _children(other, Sum))

    def __mul__(self, other: "KernelSpec") -> "KernelSpec":
        return Product(_children(self, Product) + _children(other, Product))

    def program(self) -> np.ndarray:
        ops: List[int] = []
        self._emit(ops)
        return np.asarray(ops, dtype=np.int32)

    def get_params(self) -> np.ndarray:
        out: List[float] = []
        self._params(out)
        return np.asarray(out, dtype=np.float64)

    def num_params(self) -> int:
        return len(self.get_params())

    def set_params(self, params: Sequence[float]) -> None:
        rest = self._set(list(params))
        if rest:
            raise ValueError("too many parameters for this kernel")

    # subclasses implement _emit, _params, _set


def _children(k: KernelSpec, cls) -> list:
    return list(k.children) if isinstance(k, cls) else [k]


class _Leaf(KernelSpec):
    pass


class ConstantKernel(_Leaf):
    """k = c; theta = log c (src/kernel/constant.rs:12-16)."""

    def __init__(self, constant: float):
        if not (constant > 0 and math.isfinite(constant)):
            raise ValueError("constant must be positive and finite")
        self.log_c = math.log(constant)

    def _emit(self, ops):
        ops += [OP_CONST, 0]

    def _params(self, out):
        out.append(self.log_c)

    def _set(self, p):
        self.log_c = p[0]
        return p[1:]


class RbfKernel(_Leaf):
    """k = exp(-|x - x'|^2 / (2 l^2)); theta = log l (src/kernel/rbf.rs:18-20)."""

    def __init__(self, lengthscale: float):
        if not (lengthscale > 0 and math.isfinite(lengthscale)):
            raise ValueError("lengthscale must be positive and finite")
        self.log_l = math.log(lengthscale)

    def _emit(self, ops):
        ops += [OP_RBF, 0]

    def _params(self, out):
        out.append(self.log_l)

    def _set(self, p):
        self.log_l = p[0]
        return p[1:]


class RbfArdKernel(_Leaf):
    """k = exp(-1/2 sum_d (x_d - x'_d)^2 / l_d^2); theta_d = log l_d (src/kernel/rbf_ard.rs:16)."""

    def __init__(self, lengthscales: Sequence[float]):
        ls = [float(v) for v in lengthscales]
        if not ls or any(not (v > 0 and math.isfinite(v)) for v in ls):
            raise ValueError("lengthscales must be positive and finite")
        self.log_l = [math.log(v) for v in ls]

    def _emit(self, ops):
        ops += [OP_RBF_ARD, len(self.log_l)]

    def _params(self, out):
        out.extend(self.log_l)

    def _set(self, p):
        d = len(self.log_l)
        self.log_l = list(p[:d])
        return p[d:]


class PeriodicKernel(_Leaf):
    """k = exp(-2 sin^2(pi r / p) / l^2); theta = [log l, log p] (src/kernel/periodic.rs:17-23)."""

    def __init__(self, lengthscale: float, period: float):
        if not (lengthscale > 0 and period > 0):
            raise ValueError("lengthscale and period must be positive")
        self.log_l = math.log(lengthscale)
        self.log_p = math.log(period)

    def _emit(self, ops):
        ops += [OP_PERIODIC, 0]

    def _params(self, out):
        out += [self.log_l, self.log_p]

    def _set(self, p):
        self.log_l, self.log_p = p[0], p[1]
        return p[2:]


class _Node(KernelSpec):
    OP = 0

    def __init__(self, children: Sequence[KernelSpec]):
        self.children = list(children)

    def _emit(self, ops):
        for c in self.children:
            c._emit(ops)
        ops += [self.OP, len(self.children)]

    def _params(self, out):
        for c in self.children:
            c._params(out)

    def _set(self, p):
        for c in self.children:
            p = c._set(p)
        return p


class Sum(_Node):
    OP = OP_SUM


class Product(_Node):
    OP = OP_PROD
