"""Gpr.fit with Lbfgs on the GPU: the restart starts of a fit evaluated
together by the batched path (gprx_hip_exact_fit).  The reference's own fit
tests (src/gpr/tests.rs: lbfgs_knobs_affect_fit_and_refit,
lbfgs_restarts_never_end_worse_than_the_first_start; tests/fit_goldens.rs),
and the same solver run on the oracle's objective."""
import math

import numpy as np
import pytest

import gprx_amd as g
import oracle
from conftest import load_golden

pytestmark = pytest.mark.gpu


def rbf_gpr(ls, noise):
    return g.Gpr(g.RbfKernel(ls), g.GaussianLikelihood(noise))


def wavy_data():
    # src/gpr/tests.rs: wavy_data
    x = np.arange(30) / 5.0
    return x, np.sin(1.3 * x) + 0.1 * np.cos(7.0 * x)


X4 = np.array([0.0, 0.25, 0.6, 1.0])
Y4 = np.array([0.1, -0.4, 0.2, 0.8])


def test_zero_iterations_keep_the_start(gpu):
    print("[fit] zero iterations", flush=True)
    f = rbf_gpr(2.0, 0.2).with_optimizer(g.Lbfgs().with_max_iterations(0)).fit(X4, 4, 1, Y4)
    p = f.get_params()
    assert abs(p[0] - math.log(2.0)) < 1e-6 and abs(p[1] - math.log(0.2)) < 1e-6
    assert f.fit_result.iterations == 0


def test_search_moves_theta_and_lowers_nlml(gpu):
    print("[fit] search", flush=True)
    start = rbf_gpr(2.0, 0.2).factor(X4, 4, 1, Y4).neg_log_marginal_likelihood()
    f = rbf_gpr(2.0, 0.2).with_optimizer(
        g.Lbfgs().with_max_iterations(80).with_history_size(1).with_tolerance(1e-8)).fit(X4, 4, 1, Y4)
    p = f.get_params()
    assert max(abs(p[0] - math.log(2.0)), abs(p[1] - math.log(0.2))) > 1e-9
    assert f.neg_log_marginal_likelihood() < start
    assert f.neg_log_marginal_likelihood() == pytest.approx(f.fit_result.value, rel=1e-12)


def test_restarts_never_end_worse_than_the_first_start(gpu):
    x, y = wavy_data()
    plain = rbf_gpr(1.0, 0.1).fit(x, 30, 1, y).neg_log_marginal_likelihood()
    for seed in range(10):
        print(f"[fit] restarts seed {seed}", flush=True)
        f = rbf_gpr(1.0, 0.1).with_optimizer(g.Lbfgs().with_restarts(3, seed)).fit(x, 30, 1, y)
        assert f.neg_log_marginal_likelihood() <= plain + 1e-8, (seed, f.fit_result)


# The logit map of src/optimizer/logit.rs:75-230 restated for the oracle run
# (every parameter on the default positive interval).
LO, HI = 1e-5, 1e5
SPAN = math.log(HI) - math.log(LO)


def z_of(log_theta):
    u = (np.asarray(log_theta) - math.log(LO)) / SPAN
    return np.log(u / (1 - u)) * (SPAN / 4)


def log_theta_of(z):
    return math.log(LO) + SPAN / (1 + np.exp(-np.asarray(z) / (SPAN / 4)))


def dlog_dz(z):
    s = 1 / (1 + np.exp(-np.asarray(z) / (SPAN / 4)))
    return SPAN * s * (1 - s) / (SPAN / 4)


@pytest.mark.parametrize("prec,tol", [(g.Precision.FP64, 1e-6), (g.Precision.MIXED, 1e-3)])
def test_gpu_fit_matches_the_solver_on_the_oracle(gpu, prec, tol):
    """The batched GPU fit against the same solver minimizing the oracle's
    NLML through the same logit map on the host: the same optimum (FP64:
    the same point; mixed: a point as good to 1e-3)."""
    rng = np.random.default_rng(5)
    n, d = 200, 3
    x = rng.uniform(0, 1, n * d)
    y = np.sin(3 * x[:n]) + np.cos(2 * x[n:2 * n]) + 0.1 * rng.standard_normal(n)
    spec = g.ConstantKernel(1.0) * g.RbfArdKernel([0.5, 0.5, 0.5])
    ops = spec.program()
    th0 = np.concatenate([spec.get_params(), [math.log(0.05)]])

    def obj(z):
        th = log_theta_of(z)
        v, gr, _, _ = oracle.exact_value_grad(ops, th, x, n, d, y)
        return v, gr * dlog_dz(z)

    ref = g.Lbfgs().minimize_unconstrained(obj, z_of(th0))
    print(f"[fit] oracle run: {ref.iterations} iterations, nlml {ref.value}", flush=True)
    f = g.Gpr(spec, g.GaussianLikelihood(0.05)).with_precision(prec).fit(x, n, d, y)
    got = f.fit_result
    print(f"[fit] gpu run: {got.iterations} iterations, nlml {got.value}", flush=True)
    assert abs(got.value - ref.value) <= tol * max(1.0, abs(ref.value)), (got, ref)
    if prec == g.Precision.FP64:
        assert np.max(np.abs(got.params - log_theta_of(ref.params))) < 1e-4
    else:
        # f32 gradients do not resolve this NLML's flat directions: the mixed
        # fit's point is judged by the oracle's NLML there
        v, _, _, _ = oracle.exact_value_grad(ops, got.params, x, n, d, y, want_grad=False)
        assert v <= ref.value + tol * abs(ref.value), (v, ref.value)


@pytest.mark.parametrize("name", ["forrester_rbf", "sphere_rbf_ard"])
def test_fit_goldens(gpu, name):
    """tests/fit_goldens.rs: from sklearn's initial theta on standardized
    targets, the fit lands near sklearn's optimum (NLML within 15%, each
    theta within 2.5 in log) and not above it by more than 0.1%."""
    print(f"[fit] golden {name}", flush=True)
    gd = load_golden(name)
    n, d = gd["n_rows"], gd["n_cols"]
    y = np.asarray(gd["y"])
    yn = (y - y.mean()) / y.std()
    ls0 = gd["lengthscales_init"]
    spec = g.RbfArdKernel(ls0) if gd["kernel"] == "rbf_ard" else g.RbfKernel(ls0[0])
    f = g.Gpr(spec, g.GaussianLikelihood(gd["noise_variance_init"])).fit(gd["x"], n, d, yn)
    lml = -f.neg_log_marginal_likelihood()
    assert abs(lml - gd["log_marginal_likelihood"]) <= 0.15 * abs(gd["log_marginal_likelihood"])
    print(f"[fit] {name}: lml {lml} against sklearn's {gd['log_marginal_likelihood']}", flush=True)
    assert lml >= gd["log_marginal_likelihood"] - 1e-3 * abs(gd["log_marginal_likelihood"])
    p = f.get_params()
    for i, ell in enumerate(gd["lengthscales"]):
        assert abs(p[i] - math.log(ell)) <= 2.5, (i, math.exp(p[i]), ell)
    assert abs(p[-1] - math.log(gd["noise_variance"])) <= 2.5
    if gd["kernel"] == "rbf_ard" and len(gd["lengthscales"]) >= 2:
        assert p[0] < p[1]


def test_start_outside_bounds_is_an_interval_error(gpu):
    with pytest.raises(g.InvalidInterval):
        rbf_gpr(2.0, 0.2).with_bounds([0.1, 1.0, 1e-5, 1e5]).fit(X4, 4, 1, Y4)


def test_trust_region_fit_reaches_the_lbfgs_optimum(gpu):
    """TrustRegion (each point's Hessian from the GPU) and Lbfgs from the
    same start land on the same NLML."""
    rng = np.random.default_rng(12)
    n, d = 160, 2
    x = rng.uniform(0, 1, n * d)
    y = np.sin(4 * x[:n]) * np.cos(3 * x[n:]) + 0.05 * rng.standard_normal(n)
    spec = g.ConstantKernel(1.0) * g.RbfArdKernel([0.5, 0.5])
    print("[fit] trust region vs L-BFGS", flush=True)
    tr = g.Gpr(spec, g.GaussianLikelihood(0.05)).with_optimizer(g.TrustRegion()).fit(x, n, d, y)
    lb = g.Gpr(spec, g.GaussianLikelihood(0.05)).with_optimizer(
        g.Lbfgs().with_max_iterations(200)).fit(x, n, d, y)
    a, b = tr.neg_log_marginal_likelihood(), lb.neg_log_marginal_likelihood()
    print(f"[fit] trust region {a} in {tr.fit_result.iterations} iterations, L-BFGS {b}", flush=True)
    assert abs(a - b) <= 1e-7 * max(1.0, abs(b))
    assert a == pytest.approx(tr.fit_result.value, rel=1e-12)


@pytest.mark.parametrize("name", ["forrester_rbf", "sphere_rbf_ard"])
def test_trust_region_fit_goldens(gpu, name):
    gd = load_golden(name)
    n, d = gd["n_rows"], gd["n_cols"]
    y = np.asarray(gd["y"])
    yn = (y - y.mean()) / y.std()
    ls0 = gd["lengthscales_init"]
    spec = g.RbfArdKernel(ls0) if gd["kernel"] == "rbf_ard" else g.RbfKernel(ls0[0])
    f = g.Gpr(spec, g.GaussianLikelihood(gd["noise_variance_init"])).with_optimizer(
        g.TrustRegion()).fit(gd["x"], n, d, yn)
    lml = -f.neg_log_marginal_likelihood()
    print(f"[fit] trust region {name}: lml {lml} against sklearn's {gd['log_marginal_likelihood']}",
          flush=True)
    # a local method from sklearn's start: at sklearn's optimum or a better one
    assert lml >= gd["log_marginal_likelihood"] - 1e-3 * abs(gd["log_marginal_likelihood"])
    assert np.all(np.isfinite(f.get_params()))


# Two coincident points: with noise 1e-18, 1 + sigma^2 rounds to 1 and the
# factor meets an exactly-zero pivot unless jitter is added.
XD = np.array([0.0, 0.0, 0.25, 0.6, 1.0])
YD = np.array([0.1, 0.3, -0.4, 0.2, 0.8])
BOUNDS_D = [1e-5, 1e5, 1e-20, 1e5]


def test_fit_takes_the_models_jitter_policy(gpu):
    """Every evaluation of the reference's fit uses the model's jitter policy
    (src/gpr/exact_fit.rs:280-283): from a start whose factor needs jitter,
    the fit with an adaptive policy runs, and its result is the policy's
    NLML at the fitted point on the single-model path."""
    print("[fit] jitter policy", flush=True)
    pol = g.JitterPolicy.adaptive(1e-8, 10.0, 5, 1e-3)
    gp = rbf_gpr(1.0, 1e-18).with_jitter_policy(pol).with_bounds(BOUNDS_D)
    start = gp.factor(XD, 5, 1, YD)
    v0 = start.neg_log_marginal_likelihood()
    assert start.last_jitter() > 0.0  # the start itself needs jitter
    f = rbf_gpr(1.0, 1e-18).with_jitter_policy(pol).with_bounds(BOUNDS_D).with_optimizer(
        g.Lbfgs().with_max_iterations(50)).fit(XD, 5, 1, YD)
    got = f.fit_result
    print(f"[fit] start nlml {v0} fitted {got.value} after {got.iterations} iterations", flush=True)
    assert np.isfinite(got.value) and got.value <= v0
    single = rbf_gpr(1.0, 1e-18).with_jitter_policy(pol).with_bounds(BOUNDS_D).factor(XD, 5, 1, YD)
    gr = np.empty(2)
    assert single.value_and_gradient_into(got.params, gr) == pytest.approx(got.value, rel=1e-12)


@pytest.mark.parametrize("opt", ["lbfgs", "trust_region"])
def test_unfactorable_start_fails_with_its_own_error(gpu, opt):
    """With the default fixed(0) policy the start above cannot be factored:
    the fit returns that evaluation's CholeskyFailed
    (src/optimizer/mod.rs:210-214, src/optimizer/trust_region.rs:166-168),
    not a convergence failure."""
    print(f"[fit] unfactorable start, {opt}", flush=True)
    o = g.Lbfgs() if opt == "lbfgs" else g.TrustRegion()
    with pytest.raises(g.CholeskyFailed):
        rbf_gpr(1.0, 1e-18).with_bounds(BOUNDS_D).with_optimizer(o).fit(XD, 5, 1, YD)
