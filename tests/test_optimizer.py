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
