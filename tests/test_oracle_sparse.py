"""Pins the oracle's VFE (Sgpr) and SVGP restatements (CPU only).

Goldens: GPyTorch fixtures sgpr_*.json / svgp_*.json at the reference's 1e-8
(tests/sparse_gpytorch.rs:12), rbf and rbf_ard entries.  The VFE gradient has
no golden: it is pinned by central differences and by Z = X reducing to the
exact model (src/sgpr/tests.rs:705-832).  SVGP: prior-q goldens, central
differences of every parameter group, and the Titsias-optimal q reducing to
the VFE bound (src/svgp/tests.rs:117-227).
"""
import math

import numpy as np
import pytest

import gprx_amd as g
import oracle
from conftest import assert_close, load_golden

SPARSE = ["sgpr_forrester", "sgpr_sphere"]


def _spec(k):
    if k["name"] == "rbf":
        return g.RbfKernel(k["lengthscales"][0])
    if k["name"] == "rbf_ard":
        return g.RbfArdKernel(k["lengthscales"])
    return None


@pytest.mark.parametrize("name", SPARSE)
def test_vfe_goldens(name):
    gd = load_golden(name)
    n, d, m, q = gd["n_rows"], gd["n_cols"], gd["m"], gd["xs_n_rows"]
    done = 0
    for k in gd["kernels"]:
        spec = _spec(k)
        if spec is None:
            continue  # Matern / White leaves are outside this backend
        th = np.concatenate([spec.get_params(), [math.log(gd["noise_variance"])]])
        val, _ = oracle.vfe_value_grad(spec.program(), th, gd["x"], n, d, gd["y"], gd["z"], m, want_grad=False)
        assert_close(val, k["neg_log_marginal_likelihood"], 1e-8, f"{name}/{k['name']} bound")
        mean, lat, obs = oracle.vfe_predict(spec.program(), th, gd["x"], n, d, gd["y"], gd["z"], m, gd["xs"], q)
        assert_close(mean, k["mean"], 1e-8, "mean")
        assert_close(lat, k["latent_variance"], 1e-8, "latent")
        assert_close(obs, k["observation_variance"], 1e-8, "obs")
        done += 1
    assert done == 2


@pytest.mark.parametrize("name", ["svgp_forrester", "svgp_sphere"])
def test_svgp_prior_goldens(name):
    gd = load_golden(name)
    n, d, m, q = gd["n_rows"], gd["n_cols"], gd["m"], gd["xs_n_rows"]
    for k in gd["kernels"]:
        spec = _spec(k)
        if spec is None:
            continue
        th = np.concatenate([spec.get_params(), [math.log(gd["noise_variance"])]])
        params = np.concatenate([th, np.zeros(m), prior_packed(m)])
        val, _ = oracle.svgp_value_grad(spec.program(), params, gd["x"], n, d, gd["y"], gd["z"], m, np.arange(n))
        assert_close(val, k["neg_elbo"], 1e-8, f"{name}/{k['name']} -elbo")
        mean, lat, obs = oracle.svgp_predict(spec.program(), params, d, gd["z"], m, gd["xs"], q)
        assert_close(mean, k["mean"], 1e-8, "mean")
        assert_close(lat, k["latent_variance"], 1e-8, "latent")
        assert_close(obs, k["observation_variance"], 1e-8, "obs")


def prior_packed(m):
    out = []
    for j in range(m):
        for i in range(j, m):
            out.append(1.0 if i == j else 0.0)
    return np.array(out)


def pack_lower(Lq):
    m = Lq.shape[0]
    return np.array([Lq[i, j] for j in range(m) for i in range(j, m)])


def _sparse_problem(n=60, m=9, d=2, seed=0):
    rng = np.random.default_rng(seed)
    x = rng.random(n * d)
    z = rng.random(m * d)
    y = np.sin(4 * x[:n]) + 0.3 * np.cos(3 * x[n:2 * n]) + 0.05 * rng.standard_normal(n)
    spec = g.ConstantKernel(1.3) * g.RbfArdKernel([0.4, 0.6]) + g.PeriodicKernel(0.9, 1.4)
    th = np.concatenate([spec.get_params(), [math.log(0.02)]])
    return x, y, z, spec, th, n, m, d


def test_vfe_gradient_central_differences():
    x, y, z, spec, th, n, m, d = _sparse_problem()
    ops = spec.program()
    _, grad = oracle.vfe_value_grad(ops, th, x, n, d, y, z, m)
    h = 1e-5
    for i in range(th.size):
        tp, tm = th.copy(), th.copy()
        tp[i] += h
        tm[i] -= h
        fd = (oracle.vfe_value_grad(ops, tp, x, n, d, y, z, m, False)[0]
              - oracle.vfe_value_grad(ops, tm, x, n, d, y, z, m, False)[0]) / (2 * h)
        assert abs(fd - grad[i]) <= 1e-6 * max(1.0, abs(fd)), (i, fd, grad[i])


def test_vfe_at_z_equals_x_is_exact():
    rng = np.random.default_rng(4)
    n, d = 25, 2
    x = rng.random(n * d) * 4.0
    y = rng.standard_normal(n)
    spec = g.ConstantKernel(1.1) * g.RbfArdKernel([0.5, 0.8])
    th = np.concatenate([spec.get_params(), [math.log(0.1)]])
    v1, g1 = oracle.vfe_value_grad(spec.program(), th, x, n, d, y, x, n)
    v2, g2, _, _ = oracle.exact_value_grad(spec.program(), th, x, n, d, y)
    assert_close(v1, v2, 1e-8, "value")
    assert_close(g1, g2, 1e-6, "grad")


def test_svgp_gradient_central_differences():
    x, y, z, spec, th, n, m, d = _sparse_problem(n=40, m=6)
    rng = np.random.default_rng(9)
    Lq = np.tril(0.3 * rng.standard_normal((m, m))) + np.diag(0.5 + rng.random(m))
    params = np.concatenate([th, 0.4 * rng.standard_normal(m), pack_lower(Lq)])
    batch = rng.permutation(n)[:17]
    ops = spec.program()
    _, grad = oracle.svgp_value_grad(ops, params, x, n, d, y, z, m, batch)
    h = 1e-6
    for i in range(params.size):
        pp, pm = params.copy(), params.copy()
        pp[i] += h
        pm[i] -= h
        fd = (oracle.svgp_value_grad(ops, pp, x, n, d, y, z, m, batch)[0]
              - oracle.svgp_value_grad(ops, pm, x, n, d, y, z, m, batch)[0]) / (2 * h)
        assert abs(fd - grad[i]) <= 2e-6 * max(1.0, abs(fd)), (i, fd, grad[i])


def test_svgp_titsias_optimal_q_equals_vfe():
    x, y, z, spec, th, n, m, d = _sparse_problem(n=50, m=8, seed=3)
    ops = spec.program()
    s2 = math.exp(th[-1])
    Kmm = oracle.cross(ops, th, z, m, z, m, d)
    L = np.linalg.cholesky(Kmm)
    A = np.linalg.solve(L, oracle.cross(ops, th, z, m, x, n, d))
    B = A @ A.T + s2 * np.eye(m)
    S = s2 * np.linalg.inv(B)
    mu = np.linalg.solve(B, A @ y)
    params = np.concatenate([th, mu, pack_lower(np.linalg.cholesky(S))])
    v_svgp, _ = oracle.svgp_value_grad(ops, params, x, n, d, y, z, m, np.arange(n))
    v_vfe, _ = oracle.vfe_value_grad(ops, th, x, n, d, y, z, m, False)
    assert_close(v_svgp, v_vfe, 1e-9, "titsias")


def test_adam_first_step_closed_form():
    z = np.zeros(3)
    gr = np.array([2.0, -0.5, 1e-3])
    m1, m2 = np.zeros(3), np.zeros(3)
    t = oracle.adam_step(z, gr, m1, m2, 0, lr=0.01)
    assert t == 1
    expect = -0.01 * gr / (np.abs(gr) + 1e-8)
    assert np.allclose(z, expect, rtol=1e-12, atol=0)
