"""Sgpr (VFE) and Svgp on the GPU against the reference goldens and the oracle.

FP64 on well-conditioned problems (kappa(K_mm) <= 1e4): bound / -ELBO and
every gradient component within 1e-9 (|a - e| / max(|e|, 1)) of the oracle.
The GPU and the oracle both use the whitened A = L^-1 K_mn formulation of
the reference, the GPU with an explicit L^-1, so on the GPyTorch goldens --
whose 16 inducing points at lengthscale 1 make K_mm nearly singular -- the
two agree to the reference's own golden tolerance, 1e-8 (gradients 1e-7).
Mixed (f32 products over the data): within 1e-4 on values and 1e-2 on
gradients.
"""
import math

import numpy as np
import pytest

import gprx_amd as g
import oracle
from conftest import assert_close, load_golden, rel_err
from test_oracle_sparse import _sparse_problem, _spec, pack_lower, prior_packed

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", ["sgpr_forrester", "sgpr_sphere"])
def test_sgpr_goldens(gpu, name):
    gd = load_golden(name)
    n, d, m, q = gd["n_rows"], gd["n_cols"], gd["m"], gd["xs_n_rows"]
    for k in gd["kernels"]:
        spec = _spec(k)
        if spec is None:
            continue
        print(f"[sgpr golden] {name} {k['name']}", flush=True)
        th = np.concatenate([spec.get_params(), [math.log(gd["noise_variance"])]])
        f = g.Sgpr(spec, g.GaussianLikelihood(gd["noise_variance"])).prepare(gd["x"], n, d, gd["y"], gd["z"], m)
        grad = np.empty(th.size)
        val = f.value_and_gradient_into(th, grad)
        assert_close(val, k["neg_log_marginal_likelihood"], 1e-8, "bound")
        ov, og = oracle.vfe_value_grad(spec.program(), th, gd["x"], n, d, gd["y"], gd["z"], m)
        assert_close(val, ov, 1e-9, "bound vs oracle")
        assert_close(grad, og, 1e-7, "grad vs oracle")
        p = f.predict(gd["xs"], q, d)
        lat = f.predict_with(gd["xs"], q, d, g.VarianceKind.LATENT)
        assert_close(p.mean, k["mean"], 1e-8, "mean")
        assert_close(p.variance, k["observation_variance"], 1e-8, "obs")
        assert_close(lat.variance, k["latent_variance"], 1e-8, "latent")


def well_conditioned(n, m, ls, seed, d=2):
    """PSD sum of short-lengthscale leaves; kappa(K_mm) about 3e3 - 9e3 at the sizes below."""
    rng = np.random.default_rng(seed)
    x = rng.random(n * d)
    z = rng.random(m * d)
    y = np.sin(6 * x[:n]) + 0.3 * np.cos(5 * x[n:2 * n]) + 0.05 * rng.standard_normal(n)
    spec = g.ConstantKernel(1.3) * g.RbfArdKernel([ls, 1.3 * ls]) + g.ConstantKernel(0.5) * g.RbfKernel(1.5 * ls)
    th = np.concatenate([spec.get_params(), [math.log(0.02)]])
    return x, y, z, spec, th, n, m, d


# kappa(K_mm): 2.7e3, 8.8e3, 4.3e3
SIZES = [(1000, 50, 0.08, 1), (3000, 130, 0.05, 2), (6000, 260, 0.03, 4)]


def test_sgpr_small_composite_vs_oracle(gpu):
    x, y, z, spec, th, n, m, d = _sparse_problem(n=60, m=9, d=2, seed=60)
    f = g.Sgpr(spec, g.GaussianLikelihood(0.02)).prepare(x, n, d, y, z, m)
    grad = np.empty(th.size)
    val = f.value_and_gradient_into(th, grad)
    ov, og = oracle.vfe_value_grad(spec.program(), th, x, n, d, y, z, m)
    assert_close(val, ov, 1e-10, "bound")
    assert_close(grad, og, 1e-9, "grad")


@pytest.mark.parametrize("n,m,ls,seed", SIZES)
def test_sgpr_composite_vs_oracle(gpu, n, m, ls, seed):
    print(f"[sgpr] n={n} m={m}", flush=True)
    x, y, z, spec, th, n, m, d = well_conditioned(n, m, ls, seed)
    for prec, tol in ((g.Precision.FP64, 1e-9), (g.Precision.MIXED, 1e-4)):
        f = g.Sgpr(spec, g.GaussianLikelihood(0.02)).with_precision(prec).prepare(x, n, d, y, z, m)
        grad = np.empty(th.size)
        val = f.value_and_gradient_into(th, grad)
        ov, og = oracle.vfe_value_grad(spec.program(), th, x, n, d, y, z, m)
        print(f"  prec={prec} value err {rel_err(val, ov):.2e} grad err {rel_err(grad, og):.2e}")
        assert_close(val, ov, tol, "bound")
        assert_close(grad, og, tol if prec == g.Precision.FP64 else 1e-2, "grad")


def test_sgpr_six_leaf_factors_vs_oracle(gpu):
    """Six leaf factors over three product terms (the cross contraction's
    eight-slot kernel)."""
    x, y, z, _, _, n, m, d = well_conditioned(600, 40, 0.1, 12)
    spec = (g.ConstantKernel(0.8) * g.RbfArdKernel([0.1, 0.12]) + g.RbfKernel(0.15) * g.PeriodicKernel(0.3, 1.5)
            + g.ConstantKernel(0.3) * g.RbfKernel(0.1))
    th = np.concatenate([spec.get_params(), [math.log(0.02)]])
    for prec, tol in ((g.Precision.FP64, 1e-9), (g.Precision.MIXED, 1e-4)):
        f = g.Sgpr(spec, g.GaussianLikelihood(0.02)).with_precision(prec).prepare(x, n, d, y, z, m)
        grad = np.empty(th.size)
        val = f.value_and_gradient_into(th, grad)
        ov, og = oracle.vfe_value_grad(spec.program(), th, x, n, d, y, z, m)
        assert_close(val, ov, tol, f"bound prec={prec}")
        assert_close(grad, og, tol if prec == g.Precision.FP64 else 1e-2, f"grad prec={prec}")


@pytest.mark.parametrize("n,m,ls,seed", [(50, 1, 0.3, 3), (400, 128, 0.05, 7), (400, 129, 0.05, 8)])
def test_sgpr_edge_inducing_counts_vs_oracle(gpu, n, m, ls, seed):
    """One inducing point, exactly one tile of them, and one tile and one."""
    print(f"[sgpr edge] n={n} m={m}", flush=True)
    x, y, z, spec, th, n, m, d = well_conditioned(n, m, ls, seed)
    for prec, tol in ((g.Precision.FP64, 1e-9), (g.Precision.MIXED, 1e-4)):
        f = g.Sgpr(spec, g.GaussianLikelihood(0.02)).with_precision(prec).prepare(x, n, d, y, z, m)
        grad = np.empty(th.size)
        val = f.value_and_gradient_into(th, grad)
        ov, og = oracle.vfe_value_grad(spec.program(), th, x, n, d, y, z, m)
        assert_close(val, ov, tol, f"bound prec={prec}")
        assert_close(grad, og, tol if prec == g.Precision.FP64 else 1e-2, f"grad prec={prec}")


@pytest.mark.parametrize("name", ["svgp_forrester", "svgp_sphere"])
def test_svgp_prior_goldens(gpu, name):
    gd = load_golden(name)
    n, d, m, q = gd["n_rows"], gd["n_cols"], gd["m"], gd["xs_n_rows"]
    for k in gd["kernels"]:
        spec = _spec(k)
        if spec is None:
            continue
        print(f"[svgp golden] {name} {k['name']}", flush=True)
        th = np.concatenate([spec.get_params(), [math.log(gd["noise_variance"])]])
        params = np.concatenate([th, np.zeros(m), prior_packed(m)])
        f = g.Svgp(spec, g.GaussianLikelihood(gd["noise_variance"])).prepare(gd["x"], n, d, gd["y"], gd["z"], m)
        grad = np.empty(params.size)
        val = f.value_and_gradient_into(params, grad)
        assert_close(val, k["neg_elbo"], 1e-8, "-elbo")
        ov, og = oracle.svgp_value_grad(spec.program(), params, gd["x"], n, d, gd["y"], gd["z"], m, np.arange(n))
        assert_close(grad, og, 1e-5, "grad vs oracle")
        p = f.predict_with(gd["xs"], q, d, params)
        assert_close(p.mean, k["mean"], 1e-8, "mean")
        assert_close(p.variance, k["observation_variance"], 1e-8, "obs")


# kappa(K_mm): 1.4e3, 4.0e4, 3.0e3
@pytest.mark.parametrize("n,m,ls,seed,b", [(40, 6, 0.3, 4, 17), (2000, 50, 0.08, 5, 256), (6000, 260, 0.03, 9, 1000)])
def test_svgp_random_q_vs_oracle(gpu, n, m, ls, seed, b):
    print(f"[svgp] n={n} m={m} b={b}", flush=True)
    x, y, z, spec, th, n, m, d = well_conditioned(n, m, ls, seed)
    rng = np.random.default_rng(9)
    Lq = np.tril(0.1 * rng.standard_normal((m, m))) + np.diag(0.5 + rng.random(m))
    params = np.concatenate([th, 0.4 * rng.standard_normal(m), pack_lower(Lq)])
    batch = rng.permutation(n)[:b]
    ov, og = oracle.svgp_value_grad(spec.program(), params, x, n, d, y, z, m, batch)
    for prec, tol in ((g.Precision.FP64, 1e-9), (g.Precision.MIXED, 1e-4)):
        f = g.Svgp(spec, g.GaussianLikelihood(0.02)).with_precision(prec).prepare(x, n, d, y, z, m)
        grad = np.empty(params.size)
        val = f.value_and_gradient_into(params, grad, batch)
        print(f"  prec={prec} value err {rel_err(val, ov):.2e} grad err {rel_err(grad, og):.2e}")
        assert_close(val, ov, tol, "-elbo")
        assert_close(grad, og, tol if prec == g.Precision.FP64 else 1e-2, "grad")


@pytest.mark.parametrize("b", [1, 40])
def test_svgp_one_point_and_full_batch_vs_oracle(gpu, b):
    """A batch of one point and a batch of every point (n = 40, m = 6)."""
    print(f"[svgp edge] b={b}", flush=True)
    x, y, z, spec, th, n, m, d = well_conditioned(40, 6, 0.3, 4)
    rng = np.random.default_rng(11)
    Lq = np.tril(0.1 * rng.standard_normal((m, m))) + np.diag(0.5 + rng.random(m))
    params = np.concatenate([th, 0.4 * rng.standard_normal(m), pack_lower(Lq)])
    batch = rng.permutation(n)[:b]
    ov, og = oracle.svgp_value_grad(spec.program(), params, x, n, d, y, z, m, batch)
    f = g.Svgp(spec, g.GaussianLikelihood(0.02)).prepare(x, n, d, y, z, m)
    grad = np.empty(params.size)
    val = f.value_and_gradient_into(params, grad, batch)
    assert_close(val, ov, 1e-9, "-elbo")
    assert_close(grad, og, 1e-9, "grad")


def test_svgp_adam_matches_oracle_loop(gpu):
    """Three Adam steps with the reference's shuffle against the oracle-driven loop."""
    from gprx_amd.rng import SeededRng

    x, y, z, spec, th, n, m, d = _sparse_problem(n=50, m=5, seed=7)
    f = g.Svgp(spec, g.GaussianLikelihood(0.02)).prepare(x, n, d, y, z, m)
    p0 = f.params.copy()
    steps, bs, lr = 5, 16, 0.01
    pg, vals = f.adam(p0, steps=steps, batch_size=bs, lr=lr, seed=3)
    # host restatement of run_adam_fit with the oracle's value/gradient
    ln_lo, ln_hi = math.log(1e-5), math.log(1e5)
    span, sc = ln_hi - ln_lo, (ln_hi - ln_lo) / 4
    nth = th.size
    zt = np.log(((p0[:nth] - ln_lo) / span) / (1 - (p0[:nth] - ln_lo) / span)) * sc
    q = p0[nth:].copy()
    diag_idx = []
    k = m
    for j in range(m):
        diag_idx.append(k)
        k += m - j
    q[diag_idx] = np.log(q[diag_idx])
    zz = np.concatenate([zt, q])
    m1, m2 = np.zeros_like(zz), np.zeros_like(zz)
    order = np.arange(n, dtype=np.uint64)
    rng = SeededRng(3)
    start, t, ref_vals = n, 0, []
    for _ in range(steps):
        if start >= n:
            rng.shuffle(order)
            start = 0
        batch = order[start:start + bs]
        user = zz.copy()
        s = 1 / (1 + np.exp(-zz[:nth] / sc))
        user[:nth] = ln_lo + s * span
        user[nth + np.array(diag_idx)] = np.exp(zz[nth + np.array(diag_idx)])
        v, gu = oracle.svgp_value_grad(spec.program(), user, x, n, d, y, z, m, batch)
        ref_vals.append(v)
        gz = gu.copy()
        gz[:nth] *= span * s * (1 - s) / sc
        gz[nth + np.array(diag_idx)] *= user[nth + np.array(diag_idx)]
        t = oracle.adam_step(zz, gz, m1, m2, t, lr=lr)
        start += bs
    assert_close(vals, ref_vals, 1e-8, "trajectory")
