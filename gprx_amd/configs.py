"""The seeded synthetic workloads of the benchmark (one RNG stream per seed).

Draw order per configuration: x column-major (dimension-major: for k, for i),
then hyperparameters, then target noise with unit_normal, then test points.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

from .rng import SeededRng


@dataclass
class ExactWorkload:
    n: int
    d: int
    x: np.ndarray  # column-major n x d
    y: np.ndarray
    xs: np.ndarray  # column-major q x d
    q: int
    lengthscales: np.ndarray
    constant: float
    noise: float

    def kernel(self):
        from .kernels import ConstantKernel, RbfArdKernel

        return ConstantKernel(self.constant) * RbfArdKernel(self.lengthscales)

    def theta(self) -> np.ndarray:
        return np.concatenate([[math.log(self.constant)], np.log(self.lengthscales), [math.log(self.noise)]])


def exact_ard(n: int = 16384, d: int = 8, q: int = 4096, seed: int = 2) -> ExactWorkload:
    """Exact GPR, Constant(1) x RbfArd: x ~ U[0,1]^d, l_k log-uniform in [0.1, 2],
    y = sum_k sin(2 pi x_k / l_k) / d + 0.05 N(0,1), sigma^2 = 0.05^2."""
    r = SeededRng(seed)
    x = r.units(n * d)
    ls = np.array([math.exp(math.log(0.1) + r.unit() * math.log(20.0)) for _ in range(d)])
    eps = r.normals(n)
    X = x.reshape(d, n)
    y = np.sin(2 * math.pi * X / ls[:, None]).sum(axis=0) / d + 0.05 * eps
    xs = r.units(q * d)
    return ExactWorkload(n, d, x, y, xs, q, ls, 1.0, 0.05 ** 2)


def exact_rbf_small(n: int = 1024, d: int = 4, q: int = 256) -> ExactWorkload:
    """Isotropic RBF l=1, sigma^2 = 0.1, x ~ U[0,1]^4 (seed 0), test points seed 1,
    y = sin(3 x0) + cos(2 x1) + 0.5 x2 x3 + 0.1 N(0,1)."""
    r = SeededRng(0)
    x = r.units(n * d)
    eps = r.normals(n)
    X = x.reshape(d, n)
    y = np.sin(3 * X[0]) + np.cos(2 * X[1]) + 0.5 * X[2] * X[3] + 0.1 * eps
    xs = SeededRng(1).units(q * d)
    return ExactWorkload(n, d, x, y, xs, q, np.ones(1), 1.0, 0.1)


def batched_problems(k: int = 256, seed: int = 5, n_lo: int = 256, n_hi: int = 4096):
    """k independent ARD-RBF problems: n_i ~ U{n_lo..n_hi}, d_i in {2, 4, 8, 16},
    lengthscales log-U[0.1, 2], noise log-U[1e-3, 1e-1], x ~ U[0,1]^d, y from a
    256-feature random Fourier draw plus noise; one stream (seed)."""
    r = SeededRng(seed)
    out = []
    for _ in range(k):
        n = n_lo + r.up_to(n_hi - n_lo)
        d = (2, 4, 8, 16)[r.up_to(3)]
        ls = np.array([math.exp(math.log(0.1) + r.unit() * math.log(20.0)) for _ in range(d)])
        noise = math.exp(math.log(1e-3) + r.unit() * math.log(100.0))
        x = r.units(n * d)
        nf = 256
        W = r.normals(nf * d).reshape(nf, d) / ls[None, :]
        b = 2 * math.pi * r.units(nf)
        a = r.normals(nf)
        X = x.reshape(d, n).T
        f = (np.cos(X @ W.T + b[None, :]) @ a) * math.sqrt(2.0 / nf)
        y = f + math.sqrt(noise) * r.normals(n)
        out.append((n, d, x, y, ls, noise))
    return out
