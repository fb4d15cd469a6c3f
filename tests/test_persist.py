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
