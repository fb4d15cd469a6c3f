"""save / save_with_factor / load of exact models on the GPU (src/persist/,
FittedGpr::save{,_with_factor} at src/gpr/fitted.rs:810-830,
LoadedGpr::load at src/persist/mod.rs:308-310).  A loaded model must predict
as the saved one did (the parameters pass through exp and log once, so to
1e-12); the factor a save writes must be the Cholesky factor of the oracle's
K + (sigma^2 + jitter) I and alpha its solve with y."""
import json

import numpy as np
import pytest
import safetensors.numpy as stnp

import gprx_amd as g
import oracle
from gprx_amd import persist as P

pytestmark = pytest.mark.gpu


def _data(n, d, seed):
    r = np.random.default_rng(seed)
    x = r.uniform(-2.0, 2.0, size=n * d)  # column-major n x d
    cols = x.reshape(d, n)
    y = np.sin(cols.sum(axis=0)) + 0.1 * r.normal(size=n)
    xs = r.uniform(-2.0, 2.0, size=40 * d)
    return x, y, xs


def _spec(d):
    return g.ConstantKernel(1.3) * g.RbfArdKernel([0.8 + 0.2 * q for q in range(d)])


def _assert_same_predictions(a, b, xs, d, rtol):
    pa, pb = a.predict(xs, xs.size // d, d), b.predict(xs, xs.size // d, d)
    np.testing.assert_allclose(pb.mean, pa.mean, rtol=rtol, atol=rtol)
    np.testing.assert_allclose(pb.variance, pa.variance, rtol=rtol, atol=rtol)


@pytest.mark.parametrize("precision", [g.Precision.FP64, g.Precision.MIXED])
def test_save_load_round_trip(gpu, tmp_path, precision):
    n, d = 300, 3
    x, y, xs = _data(n, d, 0)
    print(f"[persist] round trip n={n} d={d} precision={precision}", flush=True)
    m = g.Gpr(_spec(d), g.GaussianLikelihood(0.05)).with_precision(precision).factor(x, n, d, y)
    m.save(tmp_path)
    cfg = json.loads((tmp_path / "config.json").read_text())
    assert cfg["format_version"] == 1 and cfg["has_factor"] is False
    assert cfg["factor_kind"] == "llt" and (cfg["n"], cfg["d"]) == (n, d)
    assert cfg.get("precision", "double") == ("mixed" if precision == g.Precision.MIXED else "double")
    # mixed residuals use the kernel evaluated again in f64 (ReevaluateKernel)
    assert cfg.get("residual") == ("reevaluate_kernel" if precision == g.Precision.MIXED else None)
    t = stnp.load_file(str(tmp_path / "model.safetensors"))
    assert set(t) == {"x", "y"}
    np.testing.assert_array_equal(t["x"].reshape(-1), x)
    np.testing.assert_array_equal(t["y"], y)
    back = g.FittedGpr.load(tmp_path)
    np.testing.assert_allclose(back.get_params(), m.get_params(), rtol=0, atol=1e-14)
    np.testing.assert_allclose(back.neg_log_marginal_likelihood(), m.neg_log_marginal_likelihood(),
                               rtol=1e-12)
    _assert_same_predictions(m, back, xs, d, 1e-10 if precision == g.Precision.MIXED else 1e-12)


def test_save_with_factor_writes_the_cholesky_factor(gpu, tmp_path):
    n, d = 260, 2
    x, y, xs = _data(n, d, 1)
    spec = _spec(d)
    noise = 0.02
    print(f"[persist] save_with_factor n={n} d={d}", flush=True)
    m = g.Gpr(spec, g.GaussianLikelihood(noise)).factor(x, n, d, y)
    m.save_with_factor(tmp_path)
    t = stnp.load_file(str(tmp_path / "model.safetensors"))
    assert t["l"].dtype == np.float64 and t["l"].shape == (n, n)
    lower = t["l"].reshape(-1).reshape(n, n).T  # column-major bytes: [col][row]
    assert np.all(np.triu(lower, 1) == 0.0)
    K = oracle.cross(spec.program(), spec.get_params(), x, n, x, n, d) + noise * np.eye(n)
    scale = np.abs(K).max()
    np.testing.assert_allclose(lower @ lower.T, K, rtol=0, atol=1e-11 * scale)
    np.testing.assert_allclose(lower, np.linalg.cholesky(K), rtol=0, atol=1e-11 * np.sqrt(scale))
    np.testing.assert_allclose(t["alpha"], np.linalg.solve(K, y), rtol=1e-9,
                               atol=1e-9 * np.abs(t["alpha"]).max())
    back = g.FittedGpr.load(tmp_path)
    _assert_same_predictions(m, back, xs, d, 1e-12)


def test_mixed_save_with_factor_writes_f32(gpu, tmp_path):
    n, d = 200, 2
    x, y, xs = _data(n, d, 2)
    m = g.Gpr(_spec(d), g.GaussianLikelihood(0.05)).with_precision(g.Precision.MIXED) \
        .factor(x, n, d, y)
    print(f"[persist] mixed save_with_factor n={n}", flush=True)
    m.save_with_factor(tmp_path)
    t = stnp.load_file(str(tmp_path / "model.safetensors"))
    assert t["l"].dtype == np.float32 and t["alpha"].dtype == np.float64
    back = g.FittedGpr.load(tmp_path)
    assert back._gpr.precision == g.Precision.MIXED
    _assert_same_predictions(m, back, xs, d, 1e-10)


def test_jitter_and_bounds_are_kept(gpu, tmp_path):
    n, d = 64, 1
    x = np.repeat(np.linspace(0.0, 1.0, n // 2), 2)  # every point twice
    y = np.cos(3.0 * x)
    pol = g.JitterPolicy.adaptive(1e-9, 10.0, 8, 1e-2)
    bounds = [1e-3, 1e3, 1e-13, 1e2]
    print("[persist] duplicated points, adaptive jitter, tiny noise", flush=True)
    m = g.Gpr(g.RbfKernel(0.3), g.GaussianLikelihood(1e-12)).with_jitter_policy(pol) \
        .with_bounds(bounds).factor(x, n, d, y)
    m.save(tmp_path)
    cfg = json.loads((tmp_path / "config.json").read_text())
    assert cfg["jitter"] == {"adaptive": {"initial": 1e-9, "multiplier": 10.0, "max_retries": 8,
                                          "max_jitter": 1e-2}}
    assert cfg.get("factor_jitter", 0.0) == m.last_jitter()
    assert cfg["kernel"]["rbf"]["lengthscale"]["lo"] == 1e-3
    back = g.FittedGpr.load(tmp_path)
    assert back.last_jitter() == m.last_jitter()
    assert back._gpr.jitter == pol and list(back._gpr.bounds) == bounds
    _assert_same_predictions(m, back, np.linspace(-0.2, 1.2, 30), d, 1e-12)


def test_online_model_saves_its_current_points(gpu, tmp_path):
    n, d = 150, 2
    x, y, xs = _data(n + 6, d, 3)
    cols = x.reshape(d, n + 6)
    m = g.Gpr(_spec(d), g.GaussianLikelihood(0.05)).factor(cols[:, :n].reshape(-1), n, d, y[:n])
    print("[persist] 6 insertions and 2 deletions, then save", flush=True)
    for i in range(n, n + 6):
        m.insert(cols[:, i], y[i])
    m.delete(3)
    m.delete(n + 2)
    keep = [i for i in range(n + 6) if i not in (3, n + 3)]
    m.save(tmp_path)
    t = stnp.load_file(str(tmp_path / "model.safetensors"))
    np.testing.assert_array_equal(t["x"].reshape(-1), cols[:, keep].reshape(-1))
    np.testing.assert_array_equal(t["y"], y[keep])
    back = g.FittedGpr.load(tmp_path)
    assert back.n == len(keep)
    _assert_same_predictions(m, back, xs, d, 1e-9)


def test_load_refuses_before_touching_the_device(gpu, tmp_path):
    n, d = 40, 1
    x, y, _ = _data(n, d, 4)
    g.Gpr(g.RbfKernel(1.0), g.GaussianLikelihood(0.1)).factor(x, n, d, y).save(tmp_path)
    cfg = json.loads((tmp_path / "config.json").read_text())
    cfg["format_version"] = 2
    (tmp_path / "config.json").write_text(json.dumps(cfg))
    with pytest.raises(g.UnsupportedPersistVersion):
        g.FittedGpr.load(tmp_path)


@pytest.mark.parametrize("with_factor", [False, True])
def test_rbf_save_load_matches_predict(gpu, tmp_path, with_factor):
    # rbf_save_load_matches_predict / rbf_save_with_factor_matches_predict
    # (src/persist/mod.rs:1166-1208): the two-point fit, saved and loaded,
    # predicts at 0.5 as before
    print(f"[persist] two-point rbf fit, with_factor={with_factor}", flush=True)
    fitted = g.Gpr(g.RbfKernel(1.0), g.GaussianLikelihood(0.1)).fit([0.0, 1.0], 2, 1, [0.0, 1.0])
    want = fitted.predict([0.5], 1, 1)
    (fitted.save_with_factor if with_factor else fitted.save)(tmp_path)
    got = g.FittedGpr.load(tmp_path).predict([0.5], 1, 1)
    np.testing.assert_allclose(got.mean, want.mean, rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(got.variance, want.variance, rtol=1e-12, atol=1e-12)


def test_failed_tensor_write_leaves_the_previous_config(gpu, tmp_path):
    # failed_tensor_write_leaves_the_previous_config (src/persist/mod.rs:1448-1483)
    def fixed_rbf(x, y):
        return g.Gpr(g.RbfKernel(1.0), g.GaussianLikelihood(0.1)).factor(x, len(x), 1, y)

    print("[persist] tensor write blocked by a directory", flush=True)
    fixed_rbf([0.0, 1.0], [0.0, 1.0]).save(tmp_path)
    before = (tmp_path / "config.json").read_bytes()
    (tmp_path / "model.safetensors").unlink()
    (tmp_path / "model.safetensors" / "blocker").mkdir(parents=True)
    with pytest.raises(P.PersistFailed):
        fixed_rbf([0.0, 1.0, 2.0], [0.0, 1.0, 0.5]).save(tmp_path)
    assert (tmp_path / "config.json").read_bytes() == before
    assert sorted(p.name for p in tmp_path.iterdir()) == ["config.json", "model.safetensors"]


# ------------------------------------------------------------------ sparse models
#
# save_sgpr / save_svgp and LoadedSgpr::load / LoadedSvgp::load
# (src/persist/sparse.rs:165-195, 547-556, 728-737) as
# tests/sparse_persist.rs checks them: a loaded model predicts as the saved
# one, from the same x, y and z to the bit.

@pytest.mark.parametrize("precision", [g.Precision.FP64, g.Precision.MIXED])
def test_sgpr_save_load_round_trip(gpu, tmp_path, precision):
    n, m, d = 700, 64, 3
    x, y, xs = _data(n, d, 3)
    z = x.reshape(d, n)[:, ::n // m][:, :m].reshape(-1).copy()  # column-major m x d
    print(f"[persist] sgpr round trip n={n} m={m} d={d} precision={precision}", flush=True)
    f = g.Sgpr(_spec(d), g.GaussianLikelihood(0.05)).with_precision(precision).factor(
        x, n, d, y, z, m)
    theta = f.get_params() + 0.05  # the parameters last evaluated are the ones saved
    want_nlml = f.value(theta)
    f.save(tmp_path)
    cfg = json.loads((tmp_path / "config.json").read_text())
    assert cfg["model"] == "sgpr" and (cfg["n"], cfg["m"], cfg["d"]) == (n, m, d)
    assert cfg.get("precision", "double") == ("mixed" if precision == g.Precision.MIXED else "double")
    t = stnp.load_file(str(tmp_path / "model.safetensors"))
    np.testing.assert_array_equal(t["x"].reshape(-1), x)
    np.testing.assert_array_equal(t["z"].reshape(-1), z)
    back = g.FittedSgpr.load(tmp_path)
    assert (back.n, back.m, back.d) == (n, m, d)
    np.testing.assert_allclose(back.get_params(), theta, rtol=0, atol=1e-14)
    # exp then log may move a parameter by an ulp, which can flip an f32
    # rounding in the mixed form
    tol = 1e-6 if precision == g.Precision.MIXED else 1e-12
    np.testing.assert_allclose(back.neg_log_marginal_likelihood(), want_nlml, rtol=tol)
    _assert_same_predictions(f, back, xs, d, tol)
    with pytest.raises(P.PersistFailed, match="LoadedSgpr::load"):
        g.FittedGpr.load(tmp_path)
    with pytest.raises(P.PersistFailed, match="LoadedSgpr::load"):
        g.FittedSvgp.load(tmp_path)


def test_svgp_save_load_round_trip(gpu, tmp_path):
    n, m, d = 900, 48, 2
    x, y, xs = _data(n, d, 4)
    z = x.reshape(d, n)[:, ::n // m][:, :m].reshape(-1).copy()
    print(f"[persist] svgp round trip n={n} m={m} d={d}", flush=True)
    f = g.Svgp(_spec(d), g.GaussianLikelihood(0.05)).prepare(x, n, d, y, z, m)
    params, _ = f.adam(steps=30, batch_size=64, lr=0.01, seed=3)
    f.save(tmp_path)
    t = stnp.load_file(str(tmp_path / "model.safetensors"))
    assert set(t) == {"x", "y", "z", "z_train", "q_mean", "q_l"}
    np.testing.assert_array_equal(t["q_mean"], params[f.theta.size:f.theta.size + m])
    assert not np.any(np.tril(t["q_l"], -1))  # [col, row] as read: row < col is zero
    back = g.FittedSvgp.load(tmp_path)
    np.testing.assert_allclose(back.params, params, rtol=0, atol=1e-14)
    qn = xs.size // d
    pa, pb = f.predict(xs, qn, d), back.predict(xs, qn, d)
    # the parameters pass through exp and log (an ulp each), which K_mm's
    # conditioning carries into the predictions at about 1e-11
    np.testing.assert_allclose(pb.mean, pa.mean, rtol=1e-10, atol=1e-10)
    np.testing.assert_allclose(pb.variance, pa.variance, rtol=1e-10, atol=1e-10)
    ga, gb = np.empty(params.size), np.empty(params.size)
    va = f.value_and_gradient_into(params, ga)
    vb = back.value_and_gradient_into(back.params, gb)
    np.testing.assert_allclose(vb, va, rtol=1e-12)
    np.testing.assert_allclose(gb, ga, rtol=1e-9, atol=1e-9 * np.abs(ga).max())
    with pytest.raises(P.PersistFailed, match="LoadedSvgp::load"):
        g.FittedSgpr.load(tmp_path)
