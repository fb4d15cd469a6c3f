"""Host side of save/load (src/persist/): the safetensors layout against the
safetensors library (the Python binding of the crate the reference writes
with, so equal bytes pin the header order, padding and offsets), the
config.json encodings of kernels and jitter policies, and the checks a load
makes before it touches the device (src/persist/mod.rs tests
format_version_is_one and unknown_format_version_is_rejected)."""
import json

import numpy as np
import pytest
import safetensors.numpy as stnp

import gprx_amd as g
from gprx_amd import persist as P


def _tensors(seed=0):
    r = np.random.default_rng(seed)
    n, d = 5, 3
    return {
        "y": (r.normal(size=n), (n,)),
        "x": (r.normal(size=n * d), (n, d)),
        "l": (r.normal(size=n * n).astype(np.float32), (n, n)),
        "alpha": (r.normal(size=n), (n,)),
    }


@pytest.mark.parametrize("with_factor", [False, True])
def test_safetensors_bytes_match_the_library(with_factor):
    t = _tensors()
    if not with_factor:
        t = {k: t[k] for k in ("x", "y")}
    ours = P.safetensors_bytes(t)
    lib = stnp.save({k: a.reshape(s) for k, (a, s) in t.items()})
    assert ours == lib
    back = stnp.load(ours)
    for k, (a, s) in t.items():
        np.testing.assert_array_equal(back[k], a.reshape(s))


def test_read_safetensors_round_trip(tmp_path):
    t = _tensors(1)
    (tmp_path / "m.safetensors").write_bytes(stnp.save({k: a.reshape(s) for k, (a, s) in t.items()}))
    got = P.read_safetensors(tmp_path / "m.safetensors")
    assert P._tensor(got, "x", (5, 3)).tolist() == t["x"][0].tolist()
    assert P._tensor(got, "l", (5, 5), ("F32",)).tolist() == t["l"][0].astype(np.float64).tolist()
    with pytest.raises(P.PersistFailed) as e:
        P._tensor(got, "x", (3, 5))
    assert e.value.kind == "tensor"
    with pytest.raises(P.PersistFailed):
        P._tensor(got, "l", (5, 5))  # F32 where F64 is required


def test_kernel_json_round_trip():
    spec = g.ConstantKernel(2.0) * g.RbfArdKernel([0.5, 3.0]) + g.PeriodicKernel(1.5, 0.25) \
        + g.RbfKernel(0.7)
    nparams = spec.num_params()
    bounds = [(1e-3 * (i + 1), 10.0 * (i + 1)) for i in range(nparams)]
    js = P.encode_kernel(spec, iter(bounds))
    # binary left-nested pairs, as `a * b + c + d` builds them
    assert set(js) == {"sum"} and set(js["sum"]["left"]) == {"sum"}
    assert js["sum"]["right"] == {"rbf": {"lengthscale": {"value": 0.7, "lo": bounds[-1][0],
                                                          "hi": bounds[-1][1]}}}
    text = json.dumps(js)
    got_bounds = []
    back = P.decode_kernel(json.loads(text), got_bounds)
    assert got_bounds == bounds
    np.testing.assert_allclose(back.get_params(), spec.get_params(), rtol=0, atol=1e-15)
    np.testing.assert_array_equal(back.program(), spec.program())


def test_jitter_json_round_trip():
    for pol in (g.JitterPolicy(), g.JitterPolicy.fixed(1e-6),
                g.JitterPolicy.adaptive(1e-8, 10.0, 5, 1e-3)):
        js = P.encode_jitter(pol)
        assert P.decode_jitter(json.loads(json.dumps(js))) == pol
    assert P.encode_jitter(g.JitterPolicy()) == {"fixed": {"jitter": 0.0}}
    assert set(P.encode_jitter(g.JitterPolicy.adaptive(1e-8, 10.0, 5, 1e-3))) == {"adaptive"}


@pytest.mark.parametrize("args", [(0.0, 10.0, 5, 1e-3), (-1e-8, 10.0, 5, 1e-3), (1e-8, 1.0, 5, 1e-3),
                                  (1e-8, 0.5, 5, 1e-3), (1e-8, 10.0, 0, 1e-3), (1e-3, 10.0, 5, 1e-4),
                                  (1e-8, 10.0, 5, float("inf")), (float("nan"), 10.0, 5, 1e-3)])
def test_adaptive_jitter_domain_matches_the_reference(args):
    """JitterPolicy::adaptive (src/policy.rs:266-304) refuses initial <= 0,
    multiplier <= 1, max_retries < 1 and max_jitter < initial or non-finite,
    and a config file holding such a policy does not load."""
    with pytest.raises(g.InvalidConfig):
        g.JitterPolicy.adaptive(*args)
    ini, mul, ret, mx = args
    body = {"adaptive": {"initial": ini, "multiplier": mul, "max_retries": ret, "max_jitter": mx}}
    with pytest.raises(P.PersistFailed) as e:
        P.decode_jitter(body)
    assert e.value.kind == "config"
    with pytest.raises(g.InvalidConfig):
        g.JitterPolicy.fixed(-1e-6)


@pytest.mark.parametrize("precision,residual", [("double", "reevaluate_kernel"), ("double", "promote_storage"),
                                                ("mixed", "reevaluate_kernel"), ("mixed", "promote_storage")])
def test_residual_field_loads_either_value(tmp_path, precision, residual):
    """The residual is read only with mixed precision (src/persist/config.rs:130-140);
    either value loads, the model being factored again in its own form."""
    gpr, *_ = P.read_model(_write(tmp_path, _config(precision=precision, residual=residual)))
    assert gpr.precision == (g.Precision.MIXED if precision == "mixed" else g.Precision.FP64)


def _config(**over):
    cfg = {"format_version": 1, "n": 2, "d": 1, "has_factor": False, "factor_kind": "llt",
           "kernel": {"rbf": {"lengthscale": {"value": 1.0, "lo": 1e-5, "hi": 1e5}}},
           "likelihood": {"noise_variance": {"value": 0.1, "lo": 1e-5, "hi": 1e5}},
           "jitter": {"fixed": {"jitter": 0.0}}, "distance_cache": "always",
           "x_unfitted": "identity", "y_unfitted": "identity", "x_transform": "identity",
           "y_transform": "identity"}
    cfg.update(over)
    return cfg


def _write(tmp_path, cfg, tensors=True):
    (tmp_path / "config.json").write_text(json.dumps(cfg))
    if tensors:
        t = {"x": (np.array([0.0, 1.0]), (2, 1)), "y": (np.array([0.0, 1.0]), (2,))}
        (tmp_path / "model.safetensors").write_bytes(P.safetensors_bytes(t))
    return tmp_path


def test_read_model_accepts_a_plain_config(tmp_path):
    gpr, x, y, n, d = P.read_model(_write(tmp_path, _config()))
    assert (n, d) == (2, 1) and x.tolist() == [0.0, 1.0] and y.tolist() == [0.0, 1.0]
    assert gpr.bounds is None and gpr.precision == g.Precision.FP64
    np.testing.assert_allclose(gpr.likelihood.noise_variance, 0.1)


def test_unknown_format_version_is_rejected(tmp_path):
    with pytest.raises(g.UnsupportedPersistVersion) as e:
        P.read_model(_write(tmp_path, _config(format_version=99)))
    assert e.value.found == 99 and e.value.supported == g.FORMAT_VERSION == 1


@pytest.mark.parametrize("over", [
    {"kernel": {"matern": {"lengthscale": {"value": 1.0, "lo": 1e-5, "hi": 1e5}, "nu": "half"}}},
    {"precision": "single"},
    {"math": "fast_approx"},
    {"x_transform": {"standardize": {"mean": [0.0], "std": [1.0]}}},
    {"factor_kind": "qr"},
    {"jitter": {"fixed": {"jitter": -1.0}}},
    {"residual": "promote"},
    {"likelihood": {"noise_variance": {"value": 0.0, "lo": 1e-5, "hi": 1e5}}},
])
def test_unsupported_configs_fail_on_the_host(tmp_path, over):
    with pytest.raises(P.PersistFailed) as e:
        P.read_model(_write(tmp_path, _config(**over)))
    assert e.value.kind == "config"


def test_missing_and_bad_tensor_files(tmp_path):
    with pytest.raises(P.PersistFailed) as e:
        P.read_model(_write(tmp_path, _config(), tensors=False))
    assert e.value.kind == "io"
    (tmp_path / "model.safetensors").write_bytes(b"")
    with pytest.raises(P.PersistFailed) as e:
        P.read_model(tmp_path)
    assert e.value.kind == "tensor"
    (tmp_path / "model.safetensors").write_bytes(
        P.safetensors_bytes({"x": (np.array([0.0, np.nan]), (2, 1)), "y": (np.zeros(2), (2,))}))
    with pytest.raises(P.PersistFailed) as e:
        P.read_model(tmp_path)
    assert e.value.kind == "tensor"
    _write(tmp_path, _config(has_factor=True))  # factor promised, not present
    with pytest.raises(P.PersistFailed) as e:
        P.read_model(tmp_path)
    assert "no tensor l" in str(e.value)


def test_leaf_and_sum_encoding_match_the_reference_literal():
    # the literal of src/persist/mod.rs:1387-1390 (an rbf leaf at the default
    # bounds, a binary sum)
    leaf = {"rbf": {"lengthscale": {"value": 1.0, "lo": 1e-5, "hi": 1e5}}}
    assert P.encode_kernel(g.RbfKernel(1.0), iter([P.DEFAULT_BOUNDS])) == leaf
    assert P.encode_kernel(g.RbfKernel(1.0) + g.RbfKernel(1.0), iter([P.DEFAULT_BOUNDS] * 2)) == \
        {"sum": {"left": leaf, "right": leaf}}


def test_deeply_nested_config_trees_are_rejected(tmp_path):
    # deeply_nested_config_trees_are_rejected_without_overflow
    # (src/persist/mod.rs:1370-1420): 300 nested sums are refused as a config
    # error naming the recursion limit; 20 load
    leaf = {"rbf": {"lengthscale": {"value": 1.0, "lo": 1e-5, "hi": 1e5}}}

    def nested(depth):
        k = leaf
        for _ in range(depth):
            k = {"sum": {"left": k, "right": leaf}}
        return k

    with pytest.raises(P.PersistFailed) as e:
        P.read_model(_write(tmp_path, _config(kernel=nested(300))))
    assert e.value.kind == "config" and "recursion limit" in str(e.value)
    gpr, *_ = P.read_model(_write(tmp_path, _config(kernel=nested(20))))
    assert gpr.kernel.num_params() == 21


@pytest.mark.parametrize("name", ["x", "y", "l", "alpha"])
def test_non_finite_stored_tensors_fail_to_load(tmp_path, name):
    # non_finite_stored_tensors_fail_to_load (src/persist/mod.rs:1223-1251):
    # the first value of each tensor set to NaN is a tensor error naming it
    t = {"x": (np.array([0.0, 1.0, 2.0]), (3, 1)), "y": (np.array([0.5, -1.0, 2.0]), (3,)),
         "l": (np.tril(np.ones((3, 3))).T.reshape(-1), (3, 3)), "alpha": (np.ones(3), (3,))}
    t[name] = (t[name][0].copy(), t[name][1])
    t[name][0].reshape(-1)[0] = np.nan
    (tmp_path / "config.json").write_text(json.dumps(_config(n=3, has_factor=True)))
    (tmp_path / "model.safetensors").write_bytes(P.safetensors_bytes(t))
    with pytest.raises(P.PersistFailed) as e:
        P.read_model(tmp_path)
    assert e.value.kind == "tensor"
    assert "non-finite" in str(e.value) and f"tensor {name} " in str(e.value)


def test_missing_directory_is_an_io_failure(tmp_path):
    # missing_directory_is_an_io_failure (src/persist/mod.rs:1681-1690)
    with pytest.raises(P.PersistFailed) as e:
        P.read_model(tmp_path / "does-not-exist")
    assert e.value.kind == "io"


# ------------------------------------------------------------------ sparse models

def _sparse_dir(tmp_path, kind="sgpr", mixed=False, seed=1):
    """A sparse directory written by save_sparse on the host, with what it holds."""
    r = np.random.default_rng(seed)
    n, m, d = 9, 4, 2
    x, y, z = r.normal(size=n * d), r.normal(size=n), r.normal(size=m * d)
    spec = g.ConstantKernel(1.3) * g.RbfArdKernel([0.7, 1.9])
    theta = np.concatenate([spec.get_params(), [np.log(0.05)]])
    q = None
    if kind == "svgp":
        q_mean = r.normal(size=m)
        packed = r.normal(size=m * (m + 1) // 2)
        k = 0
        for j in range(m):  # positive diagonal
            packed[k] = abs(packed[k]) + 0.1
            k += m - j
        q = (q_mean, packed)
    P.save_sparse(kind, spec, theta, x, y, z, n, m, d, mixed, tmp_path, q=q)
    return dict(n=n, m=m, d=d, x=x, y=y, z=z, theta=theta, q=q, spec=spec)


def _edit_config(path, **changes):
    cfg = json.loads((path / "config.json").read_text())
    for k, v in changes.items():
        if v is None:
            cfg.pop(k, None)
        else:
            cfg[k] = v
    (path / "config.json").write_text(json.dumps(cfg))


@pytest.mark.parametrize("kind", ["sgpr", "svgp"])
@pytest.mark.parametrize("mixed", [False, True])
def test_sparse_config_fields_in_the_structs_order(tmp_path, kind, mixed):
    """SparseConfig (src/persist/config.rs:534-565): fields in declaration
    order, precision left out at double, the sparse default K_mm policy."""
    _sparse_dir(tmp_path, kind, mixed)
    cfg = json.loads((tmp_path / "config.json").read_text())
    keys = ["format_version", "model", "n", "m", "d"] + (["precision"] if mixed else []) + \
        ["kernel", "likelihood", "jitter", "x_unfitted", "y_unfitted", "x_transform",
         "y_transform"]
    assert list(cfg) == keys
    assert cfg["model"] == kind and (cfg["n"], cfg["m"], cfg["d"]) == (9, 4, 2)
    assert cfg["jitter"] == {"adaptive": {"initial": 1e-8, "multiplier": 10.0,
                                          "max_retries": 5, "max_jitter": 1e-3}}
    nv = cfg["likelihood"]["noise_variance"]
    assert (nv["lo"], nv["hi"]) == (1e-5, 1e5) and nv["value"] == pytest.approx(0.05, rel=1e-15)
    assert cfg["kernel"]["product"]["left"]["constant"]["constant"]["value"] == \
        pytest.approx(1.3, rel=1e-15)
    assert "has_factor" not in cfg and "distance_cache" not in cfg


@pytest.mark.parametrize("kind", ["sgpr", "svgp"])
def test_sparse_tensors_and_read_back(tmp_path, kind):
    """x [n, d], y [n], z [m, d], z_train = z (identity transforms), and for
    Svgp q_mean [m] and q_l [m, m] column-major with zeros above the
    diagonal (src/persist/sparse.rs:139-154)."""
    w = _sparse_dir(tmp_path, kind)
    n, m, d = w["n"], w["m"], w["d"]
    t = stnp.load_file(str(tmp_path / "model.safetensors"))
    want = {"x", "y", "z", "z_train"} | ({"q_mean", "q_l"} if kind == "svgp" else set())
    assert set(t) == want
    assert t["x"].shape == (n, d) and t["z"].shape == (m, d)
    np.testing.assert_array_equal(t["x"].reshape(-1), w["x"])
    np.testing.assert_array_equal(t["z"].reshape(-1), w["z"])
    np.testing.assert_array_equal(t["z_train"], t["z"])
    kernel, noise, mixed, x, y, z, n2, m2, d2, q = P.read_sparse(tmp_path, kind)
    assert (n2, m2, d2, mixed) == (n, m, d, False)
    assert noise == pytest.approx(0.05, rel=1e-15)
    np.testing.assert_allclose(kernel.get_params(), w["spec"].get_params(), rtol=0, atol=1e-15)
    np.testing.assert_array_equal(x, w["x"])
    np.testing.assert_array_equal(y, w["y"])
    np.testing.assert_array_equal(z, w["z"])
    if kind == "svgp":
        # the bytes are column-major: element (row, col) at col * m + row
        ql = t["q_l"].reshape(-1)
        lower = np.zeros((m, m))
        k = 0
        for j in range(m):
            lower[j:, j] = w["q"][1][k:k + m - j]
            k += m - j
        np.testing.assert_array_equal(ql, lower.T.reshape(-1))
        np.testing.assert_array_equal(q[0], w["q"][0])
        np.testing.assert_array_equal(q[1], w["q"][1])
    else:
        assert q is None


def test_each_loader_rejects_the_other_models(tmp_path):
    """each_loader_rejects_the_other_models (tests/sparse_persist.rs:286-343):
    PersistFailed of kind "wrong model" whose message names the loader."""
    _sparse_dir(tmp_path / "sgpr", "sgpr")
    _sparse_dir(tmp_path / "svgp", "svgp")
    cases = [(lambda: P.read_model(tmp_path / "sgpr"), "LoadedSgpr::load"),
             (lambda: P.read_sparse(tmp_path / "sgpr", "svgp"), "LoadedSgpr::load"),
             (lambda: P.read_sparse(tmp_path / "svgp", "sgpr"), "LoadedSvgp::load"),
             (lambda: P.read_model(tmp_path / "svgp"), "LoadedSvgp::load")]
    exact = tmp_path / "exact"
    exact.mkdir()
    cfg = json.loads((tmp_path / "sgpr" / "config.json").read_text())
    for k in ("model", "m"):
        cfg.pop(k)
    cfg.update(has_factor=False, factor_kind="llt")
    (exact / "config.json").write_text(json.dumps(cfg))
    cases.append((lambda: P.read_sparse(exact, "sgpr"), "LoadedGpr::load"))
    for call, loader in cases:
        with pytest.raises(P.PersistFailed) as e:
            call()
        assert e.value.kind == "wrong model" and loader in str(e.value)


@pytest.mark.parametrize("change, what", [
    ({"model": "online_sgpr"}, "online"),
    ({"model": "gp"}, "model"),
    ({"format_version": 2}, "version"),
    ({"m": 0}, "positive"),
    ({"precision": "single"}, "precision"),
    ({"residual": "promote"}, "residual"),
    ({"math": "fast_approx"}, "math"),
    ({"x_transform": {"standardize": {}}}, "x_transform"),
    ({"distance": {"table": {}}}, "distance"),
    ({"jitter": {"fixed": {"jitter": 1e-7}}}, "K_mm"),
    ({"jitter": {"adaptive": {"initial": 1e-8, "multiplier": 10.0, "max_retries": 4,
                              "max_jitter": 1e-3}}}, "K_mm"),
])
def test_sparse_load_refuses_what_this_backend_cannot_hold(tmp_path, change, what):
    _sparse_dir(tmp_path, "sgpr")
    _edit_config(tmp_path, **change)
    with pytest.raises((P.PersistFailed, P.UnsupportedPersistVersion)) as e:
        P.read_sparse(tmp_path, "sgpr")
    assert what.lower() in str(e.value).lower()


def _rewrite_tensors(path, **arrays):
    t = stnp.load_file(str(path / "model.safetensors"))
    t.update(arrays)
    (path / "model.safetensors").write_bytes(stnp.save(t))


def test_sparse_load_checks_z_train_and_shapes(tmp_path):
    w = _sparse_dir(tmp_path, "sgpr")
    z = stnp.load_file(str(tmp_path / "model.safetensors"))["z"]
    bumped = z.copy()
    bumped[0, 0] = np.nextafter(bumped[0, 0], np.inf)
    _rewrite_tensors(tmp_path, z_train=bumped)
    with pytest.raises(P.PersistFailed, match="z_train"):
        P.read_sparse(tmp_path, "sgpr")
    _rewrite_tensors(tmp_path, z_train=z, z=z[:-1])
    with pytest.raises(P.PersistFailed, match="shape"):
        P.read_sparse(tmp_path, "sgpr")
    assert w["m"] == z.shape[0]


@pytest.mark.parametrize("fault", ["upper", "zero_diag", "negative_diag", "nan_mean"])
def test_svgp_load_checks_q(tmp_path, fault):
    """read_q (src/persist/sparse.rs:393-407): q_l lower triangular with a
    positive diagonal, q finite, else a tensor error."""
    w = _sparse_dir(tmp_path, "svgp")
    t = stnp.load_file(str(tmp_path / "model.safetensors"))
    ql, qm = t["q_l"].copy(), t["q_mean"].copy()  # ql[c, r]: column-major bytes
    if fault == "upper":
        ql[2, 1] = 0.5          # row 1 of column 2, above the diagonal
    elif fault == "zero_diag":
        ql[1, 1] = 0.0
    elif fault == "negative_diag":
        ql[3, 3] = -ql[3, 3]
    else:
        qm[0] = np.nan
    _rewrite_tensors(tmp_path, q_l=ql, q_mean=qm)
    with pytest.raises(P.PersistFailed) as e:
        P.read_sparse(tmp_path, "svgp")
    assert e.value.kind == "tensor"
    assert w["m"] == 4
