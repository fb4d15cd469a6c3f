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
    isotropic: bool = False  # RbfKernel(lengthscales[0]) alone

    def kernel(self):
        from .kernels import ConstantKernel, RbfArdKernel, RbfKernel

        if self.isotropic:
            return RbfKernel(float(self.lengthscales[0]))
        return ConstantKernel(self.constant) * RbfArdKernel(self.lengthscales)

    def theta(self) -> np.ndarray:
        if self.isotropic:
            return np.array([math.log(self.lengthscales[0]), math.log(self.noise)])
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
    return ExactWorkload(n, d, x, y, xs, q, np.ones(1), 1.0, 0.1, isotropic=True)


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


@dataclass
class SparseWorkload:
    n: int
    d: int
    m: int
    x: np.ndarray  # column-major n x d
    y: np.ndarray
    z: np.ndarray  # column-major m x d, rows of x
    kernel_fn: object
    theta0: np.ndarray  # kernel log-parameters then log noise

    def kernel(self):
        return self.kernel_fn()

    def theta(self) -> np.ndarray:
        return self.theta0.copy()

    @property
    def noise(self) -> float:
        return math.exp(self.theta0[-1])


def _subset_rows(r: SeededRng, x: np.ndarray, n: int, d: int, m: int) -> np.ndarray:
    """m distinct rows of x by a partial Fisher-Yates draw from the stream."""
    idx = np.arange(n)
    for i in range(m):
        j = i + r.up_to(n - 1 - i)
        idx[i], idx[j] = idx[j], idx[i]
    X = x.reshape(d, n)
    return np.ascontiguousarray(X[:, idx[:m]]).reshape(-1)


def sgpr_vfe(n: int = 1 << 20, d: int = 8, m: int = 1024, seed: int = 3) -> SparseWorkload:
    """Sgpr (VFE), Constant x RbfArd + Periodic: x ~ N(0, I_d), w ~ N(0, I/d),
    y = tanh(w'x) + 0.3 sin(4 x0) + 0.1 N(0,1); Z = m seeded rows of x.
    Initial parameters c = 1, l_k = 1, periodic l = 1 with period 20 (a
    shorter period makes the periodic Gram of 8-D Euclidean distances
    indefinite), sigma^2 = 0.01; kappa(K_mm) is about 1.7e4."""
    from .kernels import ConstantKernel, PeriodicKernel, RbfArdKernel

    r = SeededRng(seed)
    x = r.normals(n * d)
    w = r.normals(d) / math.sqrt(d)
    eps = r.normals(n)
    X = x.reshape(d, n)
    y = np.tanh(w @ X) + 0.3 * np.sin(4 * X[0]) + 0.1 * eps
    z = _subset_rows(r, x, n, d, m)

    def kern():
        return ConstantKernel(1.0) * RbfArdKernel([1.0] * d) + PeriodicKernel(1.0, 20.0)

    th = np.concatenate([kern().get_params(), [math.log(0.01)]])
    return SparseWorkload(n, d, m, x, y, z, kern, th)


def svgp_minibatch(n: int = 1 << 22, d: int = 16, m: int = 2048, seed: int = 4) -> SparseWorkload:
    """Svgp, RbfArd: x ~ U[-1,1]^d, y = sum_k x_k^2 / d + 0.2 cos(pi x0) +
    0.1 N(0,1); Z = m seeded rows of x.  Initial l_k = 1.5 (kappa(K_mm) about
    9.5e3 at m = 2048), sigma^2 = 0.01."""
    from .kernels import RbfArdKernel

    r = SeededRng(seed)
    x = -1.0 + 2.0 * r.units(n * d)
    eps = r.normals(n)
    X = x.reshape(d, n)
    y = (X ** 2).sum(axis=0) / d + 0.2 * np.cos(math.pi * X[0]) + 0.1 * eps
    z = _subset_rows(r, x, n, d, m)

    def kern():
        return RbfArdKernel([1.5] * d)

    th = np.concatenate([kern().get_params(), [math.log(0.01)]])
    return SparseWorkload(n, d, m, x, y, z, kern, th)
