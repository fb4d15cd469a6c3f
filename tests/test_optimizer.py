"""The L-BFGS solver of libgprx_hip.so on host objectives (no GPU): the
reference's own optimizer tests (src/optimizer/lbfgs.rs tests) and textbook
problems whose minimizers are known in closed form."""
import math

import numpy as np
import pytest

import gprx_amd as g


def test_lbfgs_minimizes_quadratic_bowl():
    # src/optimizer/lbfgs.rs: lbfgs_minimizes_quadratic_bowl
    calls = []

    def bowl(x):
        calls.append(1)
        return 0.5 * float(x @ x), x.copy()

    r = g.Lbfgs().with_max_iterations(50).minimize_unconstrained(bowl, [1.0, -0.5])
    assert abs(r.params[0]) < 1e-6 and abs(r.params[1]) < 1e-6
    assert r.value < 1e-10
    assert len(calls) > 0


def test_lbfgs_rosenbrock_reaches_the_minimizer():
    # f = (1 - a)^2 + 100 (b - a^2)^2, minimum 0 at (1, 1), from (-1.2, 1)
    def rosen(x):
        a, b = x
        f = (1 - a) ** 2 + 100 * (b - a * a) ** 2
        return f, np.array([-2 * (1 - a) - 400 * a * (b - a * a), 200 * (b - a * a)])

    r = g.Lbfgs().with_max_iterations(200).with_tolerance(1e-10).minimize_unconstrained(
        rosen, [-1.2, 1.0])
    assert np.max(np.abs(r.params - 1.0)) < 1e-6, r
    assert r.value < 1e-12
    assert r.iterations < 100


def test_lbfgs_ill_scaled_quadratic():
    # f = 1/2 sum_i c_i (x_i - i)^2 with c spanning 1..1e4: minimizer x_i = i
    c = np.logspace(0, 4, 10)
    t = np.arange(10.0)

    def q(x):
        r = x - t
        return 0.5 * float(np.sum(c * r * r)), c * r

    r = g.Lbfgs().with_max_iterations(200).with_tolerance(1e-9).minimize_unconstrained(
        q, np.zeros(10))
    assert np.max(np.abs(r.params - t)) < 1e-8, r
    assert r.iterations < 100


def test_zero_iterations_return_the_start():
    # the first half of src/gpr/tests.rs: lbfgs_knobs_affect_fit_and_refit
    r = g.Lbfgs().with_max_iterations(0).minimize_unconstrained(
        lambda x: (float(x @ x), 2 * x), [0.3, -0.7])
    assert list(r.params) == [0.3, -0.7]
    assert r.iterations == 0
    assert r.value == pytest.approx(0.58)


def test_history_size_one_still_converges():
    r = g.Lbfgs().with_history_size(1).with_max_iterations(100).minimize_unconstrained(
        lambda x: (0.5 * float(x @ x) + x[0], x + np.array([1.0, 0.0])), [2.0, 3.0])
    assert np.max(np.abs(r.params - np.array([-1.0, 0.0]))) < 1e-6


def test_failed_points_are_barriers_not_errors():
    # a point where the objective fails costs 1e300 (src/optimizer/adapter.rs:116-150):
    # f = x^2 fails for x < -0.5; from 2 the minimizer 0 is still found
    def f(x):
        if x[0] < -0.5:
            raise ValueError("outside")
        return float(x[0] ** 2), 2 * x

    r = g.Lbfgs().minimize_unconstrained(f, [2.0])
    assert abs(r.params[0]) < 1e-6


def test_objective_that_always_fails_does_not_converge():
    def bad(x):
        raise ValueError("never")

    with pytest.raises(g.OptimizationNotConverged):
        g.Lbfgs().minimize_unconstrained(bad, [1.0, 2.0])


def test_invalid_settings_are_config_errors():
    with pytest.raises(g.InvalidConfig):
        g.Lbfgs().with_tolerance(-1.0)
    with pytest.raises(g.InvalidConfig):
        g.Lbfgs().with_tolerance(math.inf)
    with pytest.raises(g.InvalidConfig):
        g.Lbfgs().with_history_size(0)


def test_trust_region_quadratic_bowl_in_one_step():
    # a strictly convex quadratic inside the region: Newton's step, then |g| = 0
    A = np.array([[3.0, 1.0], [1.0, 2.0]])
    b = np.array([1.0, -1.0])

    def q(x):
        return 0.5 * x @ A @ x - b @ x, A @ x - b, A

    r = g.TrustRegion().minimize_unconstrained(q, [0.2, 0.1])
    assert np.max(np.abs(r.params - np.linalg.solve(A, b))) < 1e-12
    assert r.iterations <= 2


def test_trust_region_rosenbrock_with_its_hessian():
    def rosen(x):
        a, bb = x
        f = (1 - a) ** 2 + 100 * (bb - a * a) ** 2
        gr = np.array([-2 * (1 - a) - 400 * a * (bb - a * a), 200 * (bb - a * a)])
        H = np.array([[2 - 400 * (bb - a * a) + 800 * a * a, -400 * a], [-400 * a, 200.0]])
        return f, gr, H

    r = g.TrustRegion().with_tolerance(1e-10).minimize_unconstrained(rosen, [-1.2, 1.0])
    assert np.max(np.abs(r.params - 1.0)) < 1e-8, r
    assert r.iterations < 60


def test_trust_region_leaves_a_saddle_along_negative_curvature():
    # f = x0^2 - x1^2 + x1^4 / 4: a saddle at 0, minima at x1 = +-sqrt(2)
    def f(x):
        v = x[0] ** 2 - x[1] ** 2 + x[1] ** 4 / 4
        gr = np.array([2 * x[0], -2 * x[1] + x[1] ** 3])
        H = np.array([[2.0, 0.0], [0.0, -2.0 + 3 * x[1] ** 2]])
        return v, gr, H

    r = g.TrustRegion().minimize_unconstrained(f, [0.5, 1e-3])
    assert abs(abs(r.params[1]) - math.sqrt(2)) < 1e-8 and abs(r.params[0]) < 1e-8


def test_trust_region_zero_iterations_and_bad_settings():
    r = g.TrustRegion().with_max_iterations(0).minimize_unconstrained(
        lambda x: (float(x @ x), 2 * x, 2 * np.eye(2)), [0.3, -0.7])
    assert list(r.params) == [0.3, -0.7] and r.iterations == 0
    with pytest.raises(g.InvalidConfig):
        g.TrustRegion().with_radii(2.0, 1.0)
    with pytest.raises(g.InvalidConfig):
        g.TrustRegion().with_tolerance(-1.0)
