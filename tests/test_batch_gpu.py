"""Batched exact GPR on the GPU: every problem of a batch of uneven sizes
matches the oracle (FP64: 1e-10 on the value, 1e-9 on the gradient; mixed:
1e-4 on the value and 1e-2 on the gradient, the f32 factor's kappa * u bound) and a repeated evaluation
(graph replay with new parameters) matches a fresh one."""
import math

import numpy as np
import pytest

import gprx_amd as g
import oracle
from conftest import assert_close, duplicate_start_problem, rel_err
from gprx_amd.configs import batched_problems

pytestmark = pytest.mark.gpu


def _problems(k=6, seed=11):
    probs = batched_problems(k=k, seed=seed, n_lo=100, n_hi=700)
    out = []
    for n, d, x, y, ls, noise in probs:
        spec = g.ConstantKernel(1.0) * g.RbfArdKernel(ls)
        out.append((spec, g.GaussianLikelihood(noise), x, n, d, y))
    return out


@pytest.mark.parametrize("prec,tol", [(g.Precision.FP64, 1e-10), (g.Precision.MIXED, 1e-4)])
def test_batch_vs_oracle(gpu, prec, tol):
    probs = _problems()
    print(f"[batch] k={len(probs)} sizes={[p[3] for p in probs]} prec={prec}", flush=True)
    b = g.GprBatch(probs, precision=prec)
    thetas = [np.concatenate([p[0].get_params(), [p[1].log_param()]]) for p in probs]
    for rep in range(2):
        th = [t + 0.05 * rep for t in thetas]
        nlml, grads, errs = b.value_and_gradient(th)
        for i, (spec, lik, x, n, d, y) in enumerate(probs):
            assert errs[i] is None, errs[i]
            ov, og, _, _ = oracle.exact_value_grad(spec.program(), th[i], x, n, d, y)
            assert_close(nlml[i], ov, tol, f"nlml {i} rep {rep}")
            assert_close(grads[i], og, 1e-2 if prec else 1e-9, f"grad {i} rep {rep}")


EDGE = [(1, 1), (2, 3), (127, 5), (128, 8), (129, 16), (300, 2), (256, 1)]


@pytest.mark.parametrize("prec,tol", [(g.Precision.FP64, 1e-10), (g.Precision.MIXED, 1e-4)])
def test_batch_edge_sizes_vs_oracle(gpu, prec, tol):
    """One and two points, a tile less one, exactly one and two tiles, one
    tile and a point, beside a larger problem; every padded dimension group
    (d = 1, 2, 3, 5, 8, 16)."""
    rng = np.random.default_rng(21)
    probs = []
    for n, d in EDGE:
        x = rng.random(n * d)
        y = np.sin(4 * x[:n]) + 0.1 * rng.standard_normal(n)
        ls = list(0.4 + 0.6 * rng.random(d))
        probs.append((g.ConstantKernel(1.2) * g.RbfArdKernel(ls), g.GaussianLikelihood(0.02), x, n, d, y))
    print(f"[batch edge] sizes={EDGE} prec={prec}", flush=True)
    b = g.GprBatch(probs, precision=prec)
    thetas = [np.concatenate([p[0].get_params(), [p[1].log_param()]]) for p in probs]
    nlml, grads, errs = b.value_and_gradient(thetas)
    for i, (spec, lik, x, n, d, y) in enumerate(probs):
        assert errs[i] is None, errs[i]
        ov, og, _, _ = oracle.exact_value_grad(spec.program(), thetas[i], x, n, d, y)
        assert_close(nlml[i], ov, tol, f"nlml n={n} d={d}")
        assert_close(grads[i], og, 1e-2 if prec else 1e-9, f"grad n={n} d={d}")


def test_batch_matches_single_path(gpu):
    probs = _problems(k=3, seed=12)
    b = g.GprBatch(probs)
    thetas = [np.concatenate([p[0].get_params(), [p[1].log_param()]]) for p in probs]
    nlml, grads, _ = b.value_and_gradient(thetas)
    for i, (spec, lik, x, n, d, y) in enumerate(probs):
        m = g.Gpr(spec, lik).prepare(x, n, d, y)
        gr = np.empty(thetas[i].size)
        v = m.value_and_gradient_into(thetas[i], gr)
        assert_close(nlml[i], v, 1e-11, "nlml")
        assert_close(grads[i], gr, 1e-10, "grad")


@pytest.mark.parametrize("policy", [0, 1, 2])
@pytest.mark.parametrize("below", [0, 1 << 20])
def test_every_schedule_matches_oracle(gpu, policy, below):
    """Each launch schedule (tile order over the XCDs; 128- or 64-tiles)
    computes the same products: FP64 results match the oracle."""
    probs = _problems(k=5, seed=13)
    print(f"[batch schedule] policy={policy} below={below}", flush=True)
    b = g.GprBatch(probs)
    b.set_schedule(policy, below)
    thetas = [np.concatenate([p[0].get_params(), [p[1].log_param()]]) for p in probs]
    nlml, grads, errs = b.value_and_gradient(thetas)
    assert b.schedule()[:2] == (policy, below)
    for i, (spec, lik, x, n, d, y) in enumerate(probs):
        assert errs[i] is None, errs[i]
        ov, og, _, _ = oracle.exact_value_grad(spec.program(), thetas[i], x, n, d, y)
        assert_close(nlml[i], ov, 1e-10, f"nlml {i}")
        assert_close(grads[i], og, 1e-9, f"grad {i}")


def test_autotuner_picks_a_candidate(gpu):
    probs = _problems(k=4, seed=14)
    b = g.GprBatch(probs)
    assert b.schedule()[0] == -1
    thetas = [np.concatenate([p[0].get_params(), [p[1].log_param()]]) for p in probs]
    b.value_and_gradient(thetas)
    pol, below, ms = b.schedule()
    assert pol in (0, 1, 2) and below in (0, 256, 1024)


def test_mixed_batch_falls_back_per_problem(gpu):
    """A problem whose f32 factor fails (an exactly-zero f32 pivot) is
    evaluated alone in FP64: its results equal the single FP64 model's bit
    for bit, and the other problems keep the mixed form."""
    probs = _problems(k=3, seed=15)
    spec, noise, x, n, d, y = duplicate_start_problem()
    probs.insert(1, (spec, g.GaussianLikelihood(noise), x, n, d, y))
    print(f"[batch mixed fallback] sizes={[p[3] for p in probs]}", flush=True)
    b = g.GprBatch(probs, precision=g.Precision.MIXED)
    thetas = [np.concatenate([p[0].get_params(), [p[1].log_param()]]) for p in probs]
    nlml, grads, errs = b.value_and_gradient(thetas)
    assert list(b.last_fallback()) == [0, g.Fallback.JITTER, 0, 0]
    assert all(e is None for e in errs), errs
    m = g.Gpr(spec, g.GaussianLikelihood(noise)).prepare(x, n, d, y)
    gr = np.empty(thetas[1].size)
    v = m.value_and_gradient_into(thetas[1], gr)
    assert nlml[1] == v and np.array_equal(grads[1], gr)
    for i in (0, 2, 3):
        sp, lik, xi, ni, di, yi = probs[i]
        ov, og, _, _ = oracle.exact_value_grad(sp.program(), thetas[i], xi, ni, di, yi)
        assert_close(nlml[i], ov, 1e-4, f"nlml {i}")
        assert_close(grads[i], og, 1e-2, f"grad {i}")
    # a second call reuses the kept FP64 model
    nlml2, grads2, _ = b.value_and_gradient(thetas)
    assert nlml2[1] == v and np.array_equal(grads2[1], gr)


def test_mixed_batch_refinement_fallback_per_problem(gpu):
    """The mixed batch refines alpha in its captured graph and continues the
    problems still refining under the reference's stop rules: over a sweep
    of ever smaller noise variances on one smooth RBF Gram (as in the
    single-path stall test), the ill-conditioned end stalls or fails its f32
    factor and is evaluated alone in FP64, equal to the single FP64 model bit
    for bit, while the well-conditioned start stays mixed; graph replays
    give the same results."""
    n, d = 1024, 4
    rng = np.random.default_rng(31)
    x = rng.random(n * d)
    y = np.sin(3 * x[:n]) + np.cos(2 * x[n:2 * n]) + 0.1 * rng.standard_normal(n)
    noises = [10.0 ** e for e in np.arange(-2.0, -5.51, -0.25)]
    probs = [(g.RbfKernel(0.5), g.GaussianLikelihood(s2), x, n, d, y) for s2 in noises]
    print(f"[batch mixed refinement fallback] {len(probs)} noises", flush=True)
    b = g.GprBatch(probs, precision=g.Precision.MIXED)
    thetas = [np.concatenate([p[0].get_params(), [p[1].log_param()]]) for p in probs]
    ref = []
    for (sp, lik, xi, ni, di, yi), th in zip(probs, thetas):
        m = g.Gpr(sp, lik).prepare(xi, ni, di, yi)
        gr = np.empty(th.size)
        ref.append((m.value_and_gradient_into(th, gr), gr))
        m.close()
    for rep in range(2):
        nlml, grads, errs = b.value_and_gradient(thetas)
        fb = list(b.last_fallback())
        print(f"  rep {rep} fallbacks {fb}", flush=True)
        assert all(e is None for e in errs), errs
        assert fb[0] == g.Fallback.NONE
        assert g.Fallback.REFINEMENT in fb
        for i, f in enumerate(fb):
            if f != g.Fallback.NONE:
                assert nlml[i] == ref[i][0] and np.array_equal(grads[i], ref[i][1]), i
        assert_close(nlml[0], ref[0][0], 1e-4, f"nlml 0 rep {rep}")
