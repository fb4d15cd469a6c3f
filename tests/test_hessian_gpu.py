"""The exact NLML Hessian on the GPU (gprx_hip_exact_hessian; ExactFit::
fill_hessian_from_factor, src/gpr/exact_fit.rs:484-596).  The Hessian is
the Jacobian of the gradient, and the oracle's gradient is pinned by the
goldens and by central differences of its NLML, so central differences of
the oracle's gradient pin every entry; also symmetry and agreement of the
returned NLML and gradient with value_and_gradient_into."""
import numpy as np
import pytest

import gprx_amd as g
import oracle
from conftest import rel_err

pytestmark = pytest.mark.gpu


def problem(n, d, seed, kind):
    rng = np.random.default_rng(seed)
    x = rng.uniform(0.0, 1.0, n * d)
    X = x.reshape(d, n)
    y = np.sin(4.0 * X[0]) + (0.5 * X[1] if d > 1 else 0.0) + 0.1 * rng.standard_normal(n)
    ls = list(0.4 + 0.3 * np.arange(d) / max(d, 1))
    spec = {
        "rbf": g.RbfKernel(0.7),
        "const_ard": g.ConstantKernel(1.3) * g.RbfArdKernel(ls),
        "periodic": g.PeriodicKernel(0.9, 1.7),
        "const_ard_plus_periodic": g.ConstantKernel(0.8) * g.RbfArdKernel(ls) + g.PeriodicKernel(1.1, 2.3),
        "sum_times": (g.RbfKernel(0.5) + g.PeriodicKernel(1.0, 1.5)) * (g.ConstantKernel(2.0) + g.RbfKernel(1.5)),
    }[kind]
    return x, y, spec


def fd_hessian(ops, theta, x, n, d, y, h=1e-5):
    """Central differences of the oracle's gradient, column by column."""
    p = theta.size
    H = np.empty((p, p))
    for k in range(p):
        tp, tm = theta.copy(), theta.copy()
        tp[k] += h
        tm[k] -= h
        _, gp, _, _ = oracle.exact_value_grad(ops, tp, x, n, d, y)
        _, gm, _, _ = oracle.exact_value_grad(ops, tm, x, n, d, y)
        H[:, k] = (gp - gm) / (2 * h)
    return 0.5 * (H + H.T)


@pytest.mark.parametrize("n,d,kind", [(60, 1, "rbf"), (150, 3, "const_ard"), (140, 2, "periodic"),
                                      (200, 4, "const_ard_plus_periodic"), (129, 1, "sum_times"),
                                      (300, 8, "const_ard")])
def test_hessian_matches_central_differences_of_the_oracle_gradient(gpu, n, d, kind):
    x, y, spec = problem(n, d, 7 + n, kind)
    noise = 0.05
    print(f"[hessian] n={n} d={d} {kind}", flush=True)
    m = g.Gpr(spec, g.GaussianLikelihood(noise)).prepare(x, n, d, y)
    th = m.get_params()
    p = th.size
    grad, hess = np.empty(p), np.empty(p * p)
    v = m.value_gradient_hessian_into(th, grad, hess)
    H = hess.reshape(p, p)
    assert np.array_equal(H, H.T)
    ops = spec.program()
    v_o, g_o, _, _ = oracle.exact_value_grad(ops, th, x, n, d, y)
    assert rel_err(v, v_o) < 1e-10
    assert rel_err(grad, g_o) < 1e-10
    Hfd = fd_hessian(ops, th, x, n, d, y)
    err = rel_err(H, Hfd)
    print(f"[hessian] {kind}: max rel err vs central differences {err:.2e}", flush=True)
    assert err < 1e-6, (H, Hfd)


def test_hessian_is_fp64_for_a_mixed_model_and_leaves_the_model_usable(gpu):
    n, d = 180, 3
    x, y, spec = problem(n, d, 3, "const_ard")
    f64 = g.Gpr(spec, g.GaussianLikelihood(0.05)).prepare(x, n, d, y)
    mix = g.Gpr(spec, g.GaussianLikelihood(0.05)).with_precision(g.Precision.MIXED).prepare(x, n, d, y)
    th = f64.get_params()
    p = th.size
    ga, ha, gb, hb = np.empty(p), np.empty(p * p), np.empty(p), np.empty(p * p)
    va = f64.value_gradient_hessian_into(th, ga, ha)
    vb = mix.value_gradient_hessian_into(th, gb, hb)
    assert va == vb and np.array_equal(ga, gb) and np.array_equal(ha, hb)
    # the model still evaluates and predicts at its parameters
    g2 = np.empty(p)
    assert rel_err(f64.value_and_gradient_into(th, g2), va) < 1e-12
    xs = np.ascontiguousarray(x.reshape(d, n)[:, :5].reshape(-1))  # first 5 points, column-major
    pred = f64.predict(xs, 5, d)
    assert np.all(np.isfinite(pred.mean))


def test_hessian_rejects_a_wrong_parameter_count(gpu):
    x, y, spec = problem(50, 1, 1, "rbf")
    m = g.Gpr(spec, g.GaussianLikelihood(0.1)).prepare(x, 50, 1, y)
    with pytest.raises(g.LengthMismatch):
        m.value_gradient_hessian_into(np.zeros(3), np.empty(3), np.empty(9))
