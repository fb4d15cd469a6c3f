"""The seeded workloads of the benchmark on the GPU.

Config 1 (isotropic RBF, n=1024, d=4) and config 2's generator at n=2048
against the oracle (FP64: 1e-10 on the NLML, each gradient component and the
predictions, |a - e| / max(|e|, 1)).  At the full sizes, where the oracle
would take minutes, size-independent properties: the FP64 gradient against
central differences of the GPU's own NLML (1e-5; the difference quotient at
h = 1e-4 carries about (h^2 f''' + eps kappa |f| / h) of error), the mixed
form against FP64 (1e-4 on values, 1e-2 on gradients), repeatability, and
the SVGP Adam trajectory of both forms.
"""
import math

import numpy as np
import pytest

import gprx_amd as g
import oracle
from conftest import assert_close, rel_err
from gprx_amd.configs import exact_ard, exact_rbf_small, sgpr_vfe, svgp_minibatch

pytestmark = pytest.mark.gpu


def _exact(wl, precision=g.Precision.FP64):
    return g.Gpr(wl.kernel(), g.GaussianLikelihood(wl.noise)).with_precision(precision).prepare(
        wl.x, wl.n, wl.d, wl.y)


def test_config1_vs_oracle(gpu):
    wl = exact_rbf_small()
    print(f"[config 1] n={wl.n} d={wl.d} q={wl.q}", flush=True)
    th = wl.theta()
    m = _exact(wl)
    grad = np.empty(th.size)
    val = m.value_and_gradient_into(th, grad)
    ov, og, _, _ = oracle.exact_value_grad(wl.kernel().program(), th, wl.x, wl.n, wl.d, wl.y)
    assert_close(val, ov, 1e-10, "nlml")
    assert_close(grad, og, 1e-10, "grad")
    p = m.predict(wl.xs, wl.q, wl.d)
    om, ol, oo = oracle.exact_predict(wl.kernel().program(), th, wl.x, wl.n, wl.d, wl.y, wl.xs, wl.q)
    assert_close(p.mean, om, 1e-10, "mean")
    assert_close(p.variance, oo, 1e-10, "variance")
    lat = m.predict_with(wl.xs, wl.q, wl.d, g.VarianceKind.LATENT)
    assert_close(lat.variance, ol, 1e-10, "latent variance")


def test_config2_generator_vs_oracle(gpu):
    wl = exact_ard(n=2048, q=64)
    print(f"[config 2 subset] n={wl.n} d={wl.d}", flush=True)
    th = wl.theta()
    m = _exact(wl)
    grad = np.empty(th.size)
    val = m.value_and_gradient_into(th, grad)
    ov, og, _, _ = oracle.exact_value_grad(wl.kernel().program(), th, wl.x, wl.n, wl.d, wl.y)
    assert_close(val, ov, 1e-10, "nlml")
    assert_close(grad, og, 1e-10, "grad")
    p = m.predict(wl.xs, wl.q, wl.d)
    om, _, oo = oracle.exact_predict(wl.kernel().program(), th, wl.x, wl.n, wl.d, wl.y, wl.xs, wl.q)
    assert_close(p.mean, om, 1e-10, "mean")
    assert_close(p.variance, oo, 1e-10, "variance")


def _fd_check(f_value, th, grad, comps, h=1e-4, tol=1e-5):
    for k in comps:
        e = np.zeros_like(th)
        e[k] = h
        fd = (f_value(th + e) - f_value(th - e)) / (2 * h)
        err = abs(fd - grad[k]) / max(abs(grad[k]), 1.0)
        print(f"  d/dtheta[{k}]: analytic {grad[k]:.10e} central difference {fd:.10e} err {err:.2e}")
        assert err <= tol, (k, grad[k], fd)


def test_config2_full_size(gpu):
    wl = exact_ard()
    print(f"[config 2] n={wl.n} d={wl.d}: FP64 gradient vs central differences", flush=True)
    th = wl.theta()
    m64 = _exact(wl)
    g64 = np.empty(th.size)
    v64 = m64.value_and_gradient_into(th, g64)
    again = np.empty(th.size)
    assert m64.value_and_gradient_into(th, again) == v64 and np.array_equal(again, g64)
    _fd_check(lambda t: m64.value(t), th, g64, comps=[0, 1, th.size - 1])
    m64.value_and_gradient_into(th, again)
    p64 = m64.predict(wl.xs[: 512 * wl.d], 512, wl.d)
    m64.close()
    print("[config 2] mixed vs FP64", flush=True)
    m32 = _exact(wl, g.Precision.MIXED)
    g32 = np.empty(th.size)
    v32 = m32.value_and_gradient_into(th, g32)
    print(f"  nlml err {rel_err(v32, v64):.2e} grad err {rel_err(g32, g64):.2e} refine "
          f"{m32.last_refine_steps()} fallback {m32.last_fallback()}")
    assert_close(v32, v64, 1e-4, "nlml")
    assert_close(g32, g64, 1e-2, "grad")
    p32 = m32.predict(wl.xs[: 512 * wl.d], 512, wl.d)
    assert_close(p32.mean, p64.mean, 1e-4, "mean")
    assert_close(p32.variance, p64.variance, 1e-2, "variance")


def test_config3_sgpr_full_size(gpu):
    wl = sgpr_vfe()
    print(f"[config 3] Sgpr n={wl.n} m={wl.m}: FP64 gradient vs central differences", flush=True)
    th = wl.theta()
    f64 = g.Sgpr(wl.kernel(), g.GaussianLikelihood(wl.noise)).prepare(wl.x, wl.n, wl.d, wl.y, wl.z, wl.m)
    g64 = np.empty(th.size)
    v64 = f64.value_and_gradient_into(th, g64)
    _fd_check(lambda t: f64.value(t), th, g64, comps=[0, 1, 9, th.size - 1])
    f64.close()
    f32 = g.Sgpr(wl.kernel(), g.GaussianLikelihood(wl.noise)).with_precision(g.Precision.MIXED).prepare(
        wl.x, wl.n, wl.d, wl.y, wl.z, wl.m)
    g32 = np.empty(th.size)
    v32 = f32.value_and_gradient_into(th, g32)
    print(f"  mixed: bound err {rel_err(v32, v64):.2e} grad err {rel_err(g32, g64):.2e}")
    assert_close(v32, v64, 1e-4, "bound")
    assert_close(g32, g64, 1e-2, "grad")


def test_config4_svgp_trajectory(gpu):
    wl = svgp_minibatch()
    steps, b = 50, 4096
    print(f"[config 4] Svgp n={wl.n} m={wl.m} b={b}: {steps} Adam steps, both forms", flush=True)
    out = {}
    for prec in (g.Precision.FP64, g.Precision.MIXED):
        f = g.Svgp(wl.kernel(), g.GaussianLikelihood(wl.noise)).with_precision(prec).prepare(
            wl.x, wl.n, wl.d, wl.y, wl.z, wl.m)
        p0 = f.params.copy()
        p, vals = f.adam(p0, steps=steps, batch_size=b, lr=0.01, seed=0)
        out[prec] = (p, vals)
        f.close()
    (p64, v64), (p32, v32) = out[g.Precision.FP64], out[g.Precision.MIXED]
    assert np.all(np.isfinite(v64)) and np.all(np.isfinite(v32))
    assert v64[-1] < v64[0]  # the -ELBO falls
    traj = np.abs(v32 - v64) / np.maximum(np.abs(v64), 1.0)
    print(f"  -ELBO {v64[0]:.6e} -> {v64[-1]:.6e}; mixed trajectory err max {traj.max():.2e}")
    assert traj.max() <= 1e-2
