"""Online exact GPR on the GPU: FittedGpr.insert (gprx_hip_exact_insert,
OnlineGpr::insert in src/gpr/online.rs).  The reference's libgp goldens
(tests/online_libgp.rs: from start_n points, insert the rest one by one and
check the NLML, predictive mean and observation variance after each step,
1e-8), and insertions across the 128-row padding against a model factored
from scratch at every step."""
import math

import numpy as np
import pytest

import gprx_amd as g
from conftest import load_golden, rel_err

pytestmark = pytest.mark.gpu


def point(x, n, d, i):
    return np.array([x[q * n + i] for q in range(d)])


@pytest.mark.parametrize("name", ["online_libgp_forrester", "online_libgp_sphere"])
def test_online_libgp_goldens(gpu, name):
    gd = load_golden(name)
    n, d, start = gd["n_rows"], gd["n_cols"], gd["start_n"]
    x, y = np.asarray(gd["x"]), np.asarray(gd["y"])
    xs, q = np.asarray(gd["xs"]), gd["xs_n_rows"]
    spec = g.RbfArdKernel(gd["lengthscales"]) if gd["ard"] else g.RbfKernel(gd["lengthscales"][0])
    x0 = np.concatenate([x[k * n:k * n + start] for k in range(d)])
    print(f"[online] {name}: start {start}, {n - start} insertions", flush=True)
    m = g.Gpr(spec, g.GaussianLikelihood(gd["noise_variance"])).factor(x0, start, d, y[:start])

    def check(step, nlml):
        assert step["n"] == m.n
        assert rel_err(nlml, step["neg_log_marginal_likelihood"]) < 1e-8, (step["n"], nlml)
        p = m.predict(xs, q, d)
        assert rel_err(p.mean, step["mean"]) < 1e-8, step["n"]
        assert rel_err(p.variance, step["observation_variance"]) < 1e-8, step["n"]

    check(gd["steps"][0], m.neg_log_marginal_likelihood())
    for k, i in enumerate(range(start, n), start=1):
        nlml = m.insert(point(x, n, d, i), y[i])
        check(gd["steps"][k], nlml)
    assert k + 1 == len(gd["steps"])


def test_insertions_across_the_padding_match_a_fresh_factor(gpu):
    rng = np.random.default_rng(4)
    n, d, start = 150, 3, 120
    x = rng.uniform(0, 1, n * d)
    y = np.sin(3 * x[:n]) + 0.1 * rng.standard_normal(n)
    spec = g.ConstantKernel(1.2) * g.RbfArdKernel([0.4, 0.6, 0.8])
    lik = g.GaussianLikelihood(0.05)
    xs = rng.uniform(0, 1, 4 * d)

    def cols(m_):
        return np.concatenate([x[k * n:k * n + m_] for k in range(d)])

    m = g.Gpr(spec, lik).factor(cols(start), start, d, y[:start])
    print("[online] 120 -> 150 points across the 128-row padding", flush=True)
    for i in range(start, n):
        nlml = m.insert(point(x, n, d, i), y[i])
        if i in (start, 126, 127, 128, 140, n - 1):
            ref = g.Gpr(spec, lik).factor(cols(i + 1), i + 1, d, y[:i + 1])
            assert rel_err(nlml, ref.neg_log_marginal_likelihood()) < 1e-10, i
            a, b = m.predict(xs, 4, d), ref.predict(xs, 4, d)
            assert rel_err(a.mean, b.mean) < 1e-9 and rel_err(a.variance, b.variance) < 1e-9, i
            ref.close()
    # the gradient at the grown model is the fresh model's too
    th = m.get_params()
    ga, gb = np.empty(th.size), np.empty(th.size)
    va = m.value_and_gradient_into(th, ga)
    ref = g.Gpr(spec, lik).prepare(cols(n), n, d, y)
    vb = ref.value_and_gradient_into(th, gb)
    assert rel_err(va, vb) < 1e-10 and rel_err(ga, gb) < 1e-10


def test_insert_rejects_a_wrong_dimension_and_mixed_models(gpu):
    x = np.linspace(0, 1, 20)
    y = np.sin(4 * x)
    m = g.Gpr(g.RbfKernel(0.3), g.GaussianLikelihood(0.1)).factor(x, 20, 1, y)
    with pytest.raises(g.LengthMismatch):
        m.insert([0.1, 0.2], 0.0)
    mm = g.Gpr(g.RbfKernel(0.3), g.GaussianLikelihood(0.1)).with_precision(
        g.Precision.MIXED).factor(x, 20, 1, y)
    with pytest.raises(g.Unsupported):
        mm.insert([0.5], 0.0)


def test_deletions_match_a_fresh_factor_of_the_remaining_points(gpu):
    rng = np.random.default_rng(9)
    n, d = 140, 2
    x = rng.uniform(0, 1, n * d)
    y = np.cos(5 * x[:n]) + 0.1 * rng.standard_normal(n)
    spec = g.RbfArdKernel([0.3, 0.5])
    lik = g.GaussianLikelihood(0.02)
    xs = rng.uniform(0, 1, 3 * d)
    keep = list(range(n))
    m = g.Gpr(spec, lik).factor(x, n, d, y)
    print("[online] deletions from 140 points (across the 128-row padding) and an insertion", flush=True)
    for idx in (0, 70, 11, len(keep) - 4, 5, 3, 2, 1, 0, 100, 50, 20, 7, -1):
        idx = idx if idx >= 0 else len(keep) - 1  # -1: the last point
        nlml = m.delete(idx)
        keep.pop(idx)
        k = len(keep)
        xk = np.concatenate([x[q * n + np.array(keep)] for q in range(d)])
        ref = g.Gpr(spec, lik).factor(xk, k, d, y[keep])
        assert m.n == k
        assert rel_err(nlml, ref.neg_log_marginal_likelihood()) < 1e-10, idx
        a, b = m.predict(xs, 3, d), ref.predict(xs, 3, d)
        assert rel_err(a.mean, b.mean) < 1e-9 and rel_err(a.variance, b.variance) < 1e-9, idx
        ref.close()
    # an insertion after deletions extends the refactored model
    nlml = m.insert(point(x, n, d, 70), y[70])
    keep.append(70)
    xk = np.concatenate([x[q * n + np.array(keep)] for q in range(d)])
    ref = g.Gpr(spec, lik).factor(xk, len(keep), d, y[keep])
    assert rel_err(nlml, ref.neg_log_marginal_likelihood()) < 1e-10


def test_delete_refuses_the_last_point_and_bad_indices(gpu):
    x = np.array([0.1, 0.7])
    y = np.array([0.3, -0.2])
    m = g.Gpr(g.RbfKernel(0.5), g.GaussianLikelihood(0.1)).factor(x, 2, 1, y)
    with pytest.raises(g.LengthMismatch):
        m.delete(2)
    m.delete(0)
    assert m.n == 1
    with pytest.raises(g.InsufficientData):
        m.delete(0)


def test_mixed_deletion_factors_the_remaining_points_again(gpu):
    """A mixed model's deletion forms its factor again from the remaining
    points: the same results as a fresh mixed model of them, bit for bit."""
    rng = np.random.default_rng(12)
    n, d = 300, 3
    x = rng.uniform(0, 1, n * d)
    y = np.sin(4 * x[:n]) + 0.1 * rng.standard_normal(n)
    spec = g.ConstantKernel(1.2) * g.RbfArdKernel([0.4, 0.5, 0.7])
    lik = g.GaussianLikelihood(0.05)
    m = g.Gpr(spec, lik).with_precision(g.Precision.MIXED).factor(x, n, d, y)
    keep = list(range(n))
    print("[online] mixed deletions", flush=True)
    for idx in (0, 150, 297):
        nlml = m.delete(idx)
        keep.pop(idx)
        xk = np.concatenate([x[q * n + np.array(keep)] for q in range(d)])
        ref = g.Gpr(spec, lik).with_precision(g.Precision.MIXED).factor(xk, len(keep), d, y[keep])
        assert nlml == ref.neg_log_marginal_likelihood(), idx
        ref.close()


def test_deletion_keeps_the_factors_jitter(gpu):
    """The downdate keeps the jitter the factor was formed with (the
    reference's downdated LDL^T keeps its own): after deleting one of two
    coincident points the NLML is the oracle's at noise sigma^2 + j."""
    import oracle

    rng = np.random.default_rng(13)
    n, d = 60, 2
    pts = rng.uniform(0, 1, (n, d))
    pts[1] = pts[0]
    x = pts.T.copy().ravel()
    y = np.cos(3 * pts[:, 0]) + 0.1 * rng.standard_normal(n)
    spec = g.RbfKernel(0.3)
    noise = 1e-18
    pol = g.JitterPolicy.adaptive(1e-3, 10.0, 3, 1e-1)
    m = g.Gpr(spec, g.GaussianLikelihood(noise)).with_jitter_policy(pol).factor(x, n, d, y)
    j = m.last_jitter()
    assert j > 0.0
    nlml = m.delete(1)
    keep = [i for i in range(n) if i != 1]
    xk = pts[keep].T.copy().ravel()
    th = np.array([math.log(0.3), math.log(noise + j)])
    ov, _, _, _ = oracle.exact_value_grad(spec.program(), th, xk, n - 1, d, y[keep], want_grad=False)
    assert rel_err(nlml, ov) < 1e-10
