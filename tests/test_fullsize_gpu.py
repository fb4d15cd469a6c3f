"""FP64 parity at the full sizes of the seeded workloads, against the oracle's
fixtures (tests/golden/fullsize_*.json, written by tests/golden/make_fullsize.py
from oracle/gprx_oracle.c and checked there against LAPACK):

  exact GPR n=16384, d=8: NLML, all 10 gradient components, mean, latent and
      observation variance at 256 test points, 1e-10;
  Sgpr n=2^20, m=1024: the VFE bound and all 12 gradient components, 1e-9;
  the 256-problem batch: every problem against the looped single-problem
      path, and every 16th problem against the oracle (NLML 1e-10, gradient
      1e-9: condition numbers reach 1e7 there).

Errors are |a - e| / max(|e|, 1).  The inputs are regenerated from the seeds,
and their first values are checked against the fixture before anything runs.
"""
import json
import math

import numpy as np
import pytest

import gprx_amd as g
from conftest import GOLDEN, assert_close, rel_err
from gprx_amd.configs import batched_problems, exact_ard, sgpr_vfe

pytestmark = pytest.mark.gpu


def _fixture(name):
    return json.loads((GOLDEN / name).read_text())


def test_config2_exact_n16384_vs_oracle(gpu):
    fx = _fixture("fullsize_exact_n16384.json")
    wl = exact_ard(n=fx["n"], d=fx["d"], q=fx["q"], seed=2)
    assert np.array_equal(wl.x[:8], fx["x_head"]) and np.array_equal(wl.y[:8], fx["y_head"])
    assert np.array_equal(wl.xs[:8], fx["xs_head"])
    th = wl.theta()
    assert np.array_equal(th, fx["log_theta"])
    print(f"[config 2 full size] n={wl.n} d={wl.d} q={wl.q}: FP64 vs oracle", flush=True)
    m = g.Gpr(wl.kernel(), g.GaussianLikelihood(wl.noise)).prepare(wl.x, wl.n, wl.d, wl.y)
    grad = np.empty(th.size)
    val = m.value_and_gradient_into(th, grad)
    print(f"  nlml err {rel_err(val, fx['nlml']):.2e} grad err {rel_err(grad, fx['grad']):.2e}",
          flush=True)
    assert_close(val, fx["nlml"], 1e-10, "nlml")
    assert_close(grad, fx["grad"], 1e-10, "grad")
    p = m.predict(wl.xs, wl.q, wl.d)
    lat = m.predict_with(wl.xs, wl.q, wl.d, g.VarianceKind.LATENT)
    print(f"  mean err {rel_err(p.mean, fx['mean']):.2e} variance err "
          f"{rel_err(p.variance, fx['observation_variance']):.2e} latent err "
          f"{rel_err(lat.variance, fx['latent_variance']):.2e}", flush=True)
    assert_close(p.mean, fx["mean"], 1e-10, "mean")
    assert_close(p.variance, fx["observation_variance"], 1e-10, "observation variance")
    assert_close(lat.variance, fx["latent_variance"], 1e-10, "latent variance")
    m.close()


def test_config3_sgpr_n1048576_vs_oracle(gpu):
    fx = _fixture("fullsize_sgpr_n1048576.json")
    wl = sgpr_vfe(n=fx["n"], d=fx["d"], m=fx["m"], seed=3)
    assert np.array_equal(wl.x[:8], fx["x_head"]) and np.array_equal(wl.y[:8], fx["y_head"])
    assert np.array_equal(wl.z[:8], fx["z_head"])
    th = wl.theta()
    assert np.array_equal(th, fx["log_theta"])
    print(f"[config 3 full size] Sgpr n={wl.n} m={wl.m}: FP64 vs oracle", flush=True)
    f = g.Sgpr(wl.kernel(), g.GaussianLikelihood(wl.noise)).prepare(wl.x, wl.n, wl.d, wl.y, wl.z, wl.m)
    grad = np.empty(th.size)
    val = f.value_and_gradient_into(th, grad)
    print(f"  bound err {rel_err(val, fx['value']):.2e} grad err {rel_err(grad, fx['grad']):.2e}",
          flush=True)
    assert_close(val, fx["value"], 1e-9, "bound")
    assert_close(grad, fx["grad"], 1e-9, "grad")
    f.close()


def test_config5_batch_256_vs_single_path_and_oracle(gpu):
    fx = _fixture("fullsize_batch_subset.json")
    raw = batched_problems(k=256, seed=5)
    probs = [(g.RbfArdKernel(list(ls)), g.GaussianLikelihood(noise), x, n, d, y)
             for (n, d, x, y, ls, noise) in raw]
    thetas = [np.concatenate([np.log(ls), [math.log(noise)]]) for (n, d, x, y, ls, noise) in raw]
    print(f"[config 5 full size] 256 problems, n in [{min(r[0] for r in raw)}, "
          f"{max(r[0] for r in raw)}]: FP64 batch", flush=True)
    b = g.GprBatch(probs)
    nlml, grads, errs = b.value_and_gradient(thetas)
    assert all(e is None for e in errs)
    b.close()
    worst_v = worst_g = 0.0
    for e in fx["problems"]:
        i = e["index"]
        n, d, x, y, ls, noise = raw[i]
        assert (n, d) == (e["n"], e["d"]) and np.array_equal(x[:4], e["x_head"])
        assert np.array_equal(thetas[i], e["log_theta"])
        worst_v = max(worst_v, rel_err(nlml[i], e["nlml"]))
        worst_g = max(worst_g, rel_err(grads[i], e["grad"]))
    print(f"  oracle subset ({len(fx['problems'])} problems): nlml err {worst_v:.2e} "
          f"grad err {worst_g:.2e}", flush=True)
    assert worst_v <= 1e-10 and worst_g <= 1e-9, (worst_v, worst_g)
    # every problem against the single-problem path, looped
    sv = sg = 0.0
    for i, (spec, lik, x, n, d, y) in enumerate(probs):
        m = g.Gpr(spec, lik).prepare(x, n, d, y)
        gr = np.empty(thetas[i].size)
        v = m.value_and_gradient_into(thetas[i], gr)
        m.close()
        sv = max(sv, rel_err(nlml[i], v))
        sg = max(sg, rel_err(grads[i], gr))
    print(f"  batch vs looped single path (256 problems): nlml err {sv:.2e} grad err {sg:.2e}",
          flush=True)
    assert sv <= 1e-11 and sg <= 1e-9, (sv, sg)
