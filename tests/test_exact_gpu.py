"""Exact GPR on the GPU against the reference goldens and the CPU oracle.

FP64 form: NLML, every gradient component, mean and variance within 1e-10
(|a - e| / max(|e|, 1)) of the oracle, and within the reference's 1e-8 of
its sklearn goldens.  MIXED form (f32 factor and inverse, f64 diagonal
blocks and refined alpha): within 1e-4 on the NLML and 1e-2 on the gradient,
the kappa * u_f32 bound of the f32 inverse the gradient uses.
"""
import math

import numpy as np
import pytest

import gprx_amd as g
import oracle
from conftest import (EXACT_COMPOSITE, EXACT_RBF, assert_close, duplicate_start_problem,
                      golden_problem, rel_err)

pytestmark = pytest.mark.gpu

TOL64 = 1e-10


def _fit(spec, noise, x, n, d, y, precision=g.Precision.FP64, jitter=None):
    gp = g.Gpr(spec, g.GaussianLikelihood(noise)).with_precision(precision)
    if jitter is not None:
        gp = gp.with_jitter_policy(jitter)
    return gp.prepare(x, n, d, y)


@pytest.mark.parametrize("name", EXACT_RBF + EXACT_COMPOSITE)
def test_goldens_fp64(gpu, name):
    spec, th, x, n, d, y, xs, q, gd = golden_problem(name)
    print(f"[exact golden] {name} n={n} d={d}", flush=True)
    m = _fit(spec, gd["noise_variance"], x, n, d, y)
    grad = np.empty(th.size)
    val = m.value_and_gradient_into(th, grad)
    assert_close(-val, gd["log_marginal_likelihood"], 1e-8, "lml")
    if "grad_theta" in gd:
        assert_close(-grad[: len(gd["grad_theta"])], gd["grad_theta"], 1e-8, "grad")
    if "grad_log_lengthscale" in gd:
        assert_close(-grad[0], gd["grad_log_lengthscale"], 1e-8, "grad")
    ov, og, _, _ = oracle.exact_value_grad(spec.program(), th, x, n, d, y)
    assert_close(val, ov, TOL64, "nlml vs oracle")
    assert_close(grad, og, TOL64, "grad vs oracle")
    p = m.predict(xs, q, d)
    lat = m.predict_with(xs, q, d, g.VarianceKind.LATENT)
    assert_close(p.mean, gd["mean"], 1e-8, "mean")
    assert_close(p.variance, gd["observation_variance"], 1e-8, "obs var")
    assert_close(lat.variance, gd["latent_variance"], 1e-8, "latent var")


def _problem(n, d, seed, kname):
    rng = np.random.default_rng(seed)
    x = rng.random(n * d)
    y = np.sin(3 * x[:n]) + (np.cos(2 * x[n:2 * n]) if d > 1 else 0) + 0.1 * rng.standard_normal(n)
    ls = list(0.3 + 0.7 * rng.random(d))
    specs = {
        "rbf": g.RbfKernel(0.6),
        "const_ard": g.ConstantKernel(1.3) * g.RbfArdKernel(ls),
        "const_ard_plus_periodic": g.ConstantKernel(0.8) * g.RbfArdKernel(ls) + g.PeriodicKernel(1.1, 2.3),
        "sum_times": (g.RbfKernel(0.5) + g.PeriodicKernel(1.0, 1.5)) * (g.ConstantKernel(2.0) + g.RbfKernel(1.5)),
        # six leaf factors over three product terms: the contraction's
        # eight-slot kernels below their full eight
        "six_leaves": (g.ConstantKernel(0.8) * g.RbfArdKernel(ls) + g.RbfKernel(0.7) * g.PeriodicKernel(1.1, 2.3)
                       + g.ConstantKernel(0.3) * g.RbfKernel(1.5)),
    }
    return x, y, specs[kname]


# n = 4700 pads to 37 tiles of 128: the recursive factor splits it 18 + 19 and
# the 19 again 9 + 10, so the column-split lookahead (blocks of 16 tiles and
# more, two recursion levels) runs on uneven halves at both levels (the oracle
# takes about 15 s there).
CASES = [(1, 1, "rbf"), (7, 2, "const_ard"), (128, 3, "const_ard"), (129, 4, "const_ard_plus_periodic"),
         (300, 8, "const_ard"), (257, 1, "sum_times"), (1000, 8, "const_ard"), (700, 16, "const_ard"),
         (2500, 8, "const_ard"), (4700, 8, "const_ard"), (300, 1, "six_leaves"), (200, 2, "six_leaves")]


@pytest.mark.parametrize("n,d,kname", CASES)
def test_random_vs_oracle_fp64(gpu, n, d, kname):
    print(f"[exact fp64] n={n} d={d} kernel={kname}", flush=True)
    x, y, spec = _problem(n, d, n + d, kname)
    noise = 0.05
    th = np.concatenate([spec.get_params(), [math.log(noise)]])
    m = _fit(spec, noise, x, n, d, y)
    grad = np.empty(th.size)
    val = m.value_and_gradient_into(th, grad)
    ov, og, _, _ = oracle.exact_value_grad(spec.program(), th, x, n, d, y)
    assert_close(val, ov, TOL64, "nlml")
    assert_close(grad, og, TOL64, "grad")
    q = 37
    xs = np.random.default_rng(5).random(q * d)
    p = m.predict(xs, q, d)
    om, ol, oo = oracle.exact_predict(spec.program(), th, x, n, d, y, xs, q)
    assert_close(p.mean, om, TOL64, "mean")
    assert_close(p.variance, oo, TOL64, "var")


@pytest.mark.parametrize("q", [1, 128, 129])
def test_predict_query_counts_vs_oracle(gpu, q):
    """One query point, exactly one tile of them, and one tile and one; both
    variance kinds."""
    print(f"[exact predict] q={q}", flush=True)
    n, d = 300, 3
    x, y, spec = _problem(n, d, 31, "const_ard_plus_periodic")
    th = np.concatenate([spec.get_params(), [math.log(0.05)]])
    m = _fit(spec, 0.05, x, n, d, y)
    xs = np.random.default_rng(q).random(q * d)
    p = m.predict(xs, q, d)
    lat = m.predict_with(xs, q, d, g.VarianceKind.LATENT)
    om, ol, oo = oracle.exact_predict(spec.program(), th, x, n, d, y, xs, q)
    assert_close(p.mean, om, TOL64, "mean")
    assert_close(p.variance, oo, TOL64, "obs var")
    assert_close(lat.variance, ol, TOL64, "latent var")


@pytest.mark.parametrize("n,d,kname,tol", [(300, 8, "const_ard", 1e-4), (1000, 8, "const_ard", 1e-4),
                                           (257, 1, "sum_times", 1e-4), (2500, 8, "const_ard", 1e-3)])
def test_random_vs_oracle_mixed(gpu, n, d, kname, tol):
    """The log-determinant sums n pivots that each carry an f32 error, so
    the NLML bound grows with n (2500: 1e-3 against a value of about -70)."""
    print(f"[exact mixed] n={n} d={d} kernel={kname}", flush=True)
    x, y, spec = _problem(n, d, n + d, kname)
    noise = 0.05
    th = np.concatenate([spec.get_params(), [math.log(noise)]])
    m = _fit(spec, noise, x, n, d, y, precision=g.Precision.MIXED)
    grad = np.empty(th.size)
    val = m.value_and_gradient_into(th, grad)
    ov, og, _, _ = oracle.exact_value_grad(spec.program(), th, x, n, d, y)
    print(f"  nlml err {rel_err(val, ov):.2e} grad err {rel_err(grad, og):.2e} "
          f"refine {m.last_refine_steps()} fallback {m.last_fallback()}")
    assert_close(val, ov, tol, "nlml")
    assert_close(grad, og, 1e-2, "grad")
    assert m.last_fallback() == g.Fallback.NONE


def _eval(m, th):
    grad = np.empty(th.size)
    return m.value_and_gradient_into(th, grad), grad


def test_mixed_falls_back_when_f32_factor_needs_jitter(gpu):
    """Default bound 0: the f32 factor fails at its exactly-zero pivot, so
    the evaluation is the FP64 form's at jitter 0, bit for bit."""
    spec, noise, x, n, d, y = duplicate_start_problem()
    print(f"[exact mixed jitter fallback] n={n} d={d}", flush=True)
    th = np.concatenate([spec.get_params(), [math.log(noise)]])
    m32 = _fit(spec, noise, x, n, d, y, precision=g.Precision.MIXED)
    m64 = _fit(spec, noise, x, n, d, y)
    v32, g32 = _eval(m32, th)
    v64, g64 = _eval(m64, th)
    assert m32.last_fallback() == g.Fallback.JITTER
    assert m32.last_jitter() == 0.0 and m64.last_jitter() == 0.0
    assert v32 == v64 and np.array_equal(g32, g64)
    xs = np.random.default_rng(4).random(17 * d)
    p32, p64 = m32.predict(xs, 17, d), m64.predict(xs, 17, d)
    assert np.array_equal(p32.mean, p64.mean) and np.array_equal(p32.variance, p64.variance)


def test_mixed_jitter_within_bound_stays_f32(gpu):
    """A bound at or above the policy's first offset lets the f32 factor
    take it: no fallback, and the result is the jittered system's (the
    oracle at noise sigma^2 + j)."""
    spec, noise, x, n, d, y = duplicate_start_problem()
    print(f"[exact mixed jitter within bound] n={n} d={d}", flush=True)
    th = np.concatenate([spec.get_params(), [math.log(noise)]])
    pol = g.JitterPolicy.adaptive(1e-3, 10.0, 3, 1e-2)
    m = (g.Gpr(spec, g.GaussianLikelihood(noise)).with_precision(g.Precision.MIXED)
         .with_jitter_policy(pol).with_mixed_jitter_bound(1e-2).prepare(x, n, d, y))
    v, gr = _eval(m, th)
    assert m.last_fallback() == g.Fallback.NONE
    assert m.last_jitter() == 1e-3
    thj = th.copy()
    thj[-1] = math.log(noise + 1e-3)
    ov, og, _, _ = oracle.exact_value_grad(spec.program(), thj, x, n, d, y)
    print(f"  nlml err {rel_err(v, ov):.2e} grad err {rel_err(gr[:-1], og[:-1]):.2e} "
          f"refine {m.last_refine_steps()}")
    assert_close(v, ov, 1e-4, "nlml")
    assert_close(gr[:-1], og[:-1], 1e-2, "grad")
    # d/dlog sigma^2 scales the trace by sigma^2 alone: the jitter is fixed
    assert_close(gr[-1], og[-1] * noise / (noise + 1e-3), 1e-2, "grad noise")


def test_mixed_jitter_above_bound_falls_back(gpu):
    """The policy's first offset (1e-3) is above the bound (1e-4): the f32
    failure goes to FP64 at jitter 0, where the factor holds."""
    spec, noise, x, n, d, y = duplicate_start_problem()
    print(f"[exact mixed jitter above bound] n={n} d={d}", flush=True)
    th = np.concatenate([spec.get_params(), [math.log(noise)]])
    pol = g.JitterPolicy.adaptive(1e-3, 10.0, 3, 1e-2)
    m32 = (g.Gpr(spec, g.GaussianLikelihood(noise)).with_precision(g.Precision.MIXED)
           .with_jitter_policy(pol).with_mixed_jitter_bound(1e-4).prepare(x, n, d, y))
    m64 = _fit(spec, noise, x, n, d, y, jitter=pol)
    v32, g32 = _eval(m32, th)
    v64, g64 = _eval(m64, th)
    assert m32.last_fallback() == g.Fallback.JITTER and m32.last_jitter() == 0.0
    assert v32 == v64 and np.array_equal(g32, g64)


def test_value_only_and_repeat_deterministic(gpu):
    x, y, spec = _problem(500, 4, 11, "const_ard")
    th = np.concatenate([spec.get_params(), [math.log(0.1)]])
    m = _fit(spec, 0.1, x, 500, 4, y)
    g1, g2 = np.empty(th.size), np.empty(th.size)
    v1 = m.value_and_gradient_into(th, g1)
    v2 = m.value_and_gradient_into(th, g2)
    assert v1 == v2 and np.array_equal(g1, g2)
    assert m.value(th) == v1


def test_cholesky_failure_keeps_parameters(gpu):
    x = np.array([0.0, 0.0, 1.0, 2.0])
    y = np.array([1.0, 1.0, 0.0, 0.5])
    spec = g.RbfKernel(1.0)
    good = np.array([0.0, math.log(0.1)])
    bad = np.array([0.0, math.log(1e-300)])
    m = _fit(spec, 0.1, x, 4, 1, y)
    gr = np.empty(2)
    v0 = m.value_and_gradient_into(good, gr)
    with pytest.raises(g.CholeskyFailed):
        m.value_and_gradient_into(bad, gr)
    assert np.array_equal(m.get_params(), good)
    assert m.value(good) == v0
    p = m.predict(np.array([0.5]), 1, 1)
    assert np.isfinite(p.mean[0])


def test_adaptive_jitter_matches_oracle(gpu):
    x = np.array([0.0, 0.0, 1.0, 2.0])
    y = np.array([1.0, 1.0, 0.0, 0.5])
    spec = g.RbfKernel(1.0)
    th = np.array([0.0, math.log(1e-300)])
    m = _fit(spec, 0.1, x, 4, 1, y, jitter=g.JitterPolicy.adaptive(1e-8, 10.0, 5, 1e-3))
    gr = np.empty(2)
    val = m.value_and_gradient_into(th, gr)
    ov, og, _, taken = oracle.exact_value_grad(spec.program(), th, x, 4, 1, y, jitter=(1e-8, 10.0, 5, 1e-3))
    assert m.last_jitter() == taken
    assert_close(val, ov, 1e-9, "nlml")


def test_indefinite_kernel_fails_like_the_oracle(gpu):
    # Periodic on Euclidean distance in 5-D is not positive definite here:
    # both the oracle and the GPU factor must refuse it.
    x, y, spec = _problem(257, 5, 262, "sum_times")
    th = np.concatenate([spec.get_params(), [math.log(0.05)]])
    with pytest.raises(oracle.OracleError):
        oracle.exact_value_grad(spec.program(), th, x, 257, 5, y)
    m = _fit(spec, 0.05, x, 257, 5, y)
    with pytest.raises(g.CholeskyFailed):
        m.value_and_gradient_into(th, np.empty(th.size))


def test_input_errors(gpu):
    spec = g.RbfKernel(1.0)
    with pytest.raises(g.NonFiniteInput):
        _fit(spec, 0.1, np.array([0.0, np.nan]), 2, 1, np.array([0.0, 1.0]))
    m = _fit(spec, 0.1, np.array([0.0, 1.0]), 2, 1, np.array([0.0, 1.0]))
    with pytest.raises(g.LengthMismatch):
        m.value_and_gradient_into(np.zeros(3), np.empty(3))
    with pytest.raises(g.InvalidNoiseVariance):
        m.value_and_gradient_into(np.array([0.0, np.inf]), np.empty(2))
    with pytest.raises(g.LengthMismatch):
        _fit(g.RbfArdKernel([1.0, 1.0]), 0.1, np.zeros(3), 3, 1, np.zeros(3))


def test_mixed_refinement_stall_falls_back_to_fp64(gpu):
    """An ill-conditioned noise variance makes the f32 corrections stall
    (as the reference's ill-conditioned case, src/precision/tests.rs:40-41,
    220-238).  RBF l = 0.5 on 1024 uniform points in [0, 1]^4: as the noise
    falls from 1e-2, kappa(K + sigma^2 I) grows, the corrections through the
    f32 inverse slow down and then stall (two residual ratios above 0.9, or 10
    corrections without meeting 10 n eps_64, src/precision/refine.rs:56-98)
    before the f32 factor itself fails.  At every point where refinement
    stalled the evaluation is the FP64 model's, bit for bit."""
    n, d = 1024, 4
    rng = np.random.default_rng(31)
    x = rng.random(n * d)
    y = np.sin(3 * x[:n]) + np.cos(2 * x[n:2 * n]) + 0.1 * rng.standard_normal(n)
    spec = g.RbfKernel(0.5)
    X = x.reshape(d, n).T
    K = np.exp(-((X[:, None, :] - X[None, :, :]) ** 2).sum(-1) / (2 * 0.5 ** 2))
    ev = np.linalg.eigvalsh(K)
    seen = []
    stalled = 0
    for noise in 10.0 ** np.arange(-2.0, -7.01, -0.25):
        th = np.array([math.log(0.5), math.log(noise)])
        m32 = _fit(spec, noise, x, n, d, y, precision=g.Precision.MIXED)
        v32, g32 = _eval(m32, th)
        fb, steps = m32.last_fallback(), m32.last_refine_steps()
        m32.close()
        kappa = (ev[-1] + noise) / (max(ev[0], 0.0) + noise)
        print(f"  noise {noise:.2e} kappa*eps32 {kappa * 2.0 ** -23:.2e} fallback {fb} "
              f"corrections {steps}", flush=True)
        seen.append(fb)
        if fb == g.Fallback.JITTER:
            break  # below the window: the f32 factor fails
        if fb == g.Fallback.REFINEMENT:
            stalled += 1
            m64 = _fit(spec, noise, x, n, d, y)
            v64, g64 = _eval(m64, th)
            m64.close()
            assert v32 == v64 and np.array_equal(g32, g64)
    assert seen[0] == g.Fallback.NONE
    assert stalled >= 1, seen
