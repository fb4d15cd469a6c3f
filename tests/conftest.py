import json
import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))

GOLDEN = ROOT / "tests" / "golden"


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


def load_golden(name: str) -> dict:
    return json.loads((GOLDEN / f"{name}.json").read_text())


def rel_err(a, e):
    """|a - e| / max(|e|, 1), the reference's check (tests/common/check.rs:9-19)."""
    a = np.asarray(a, dtype=np.float64)
    e = np.asarray(e, dtype=np.float64)
    return np.max(np.abs(a - e) / np.maximum(np.abs(e), 1.0)) if a.size else 0.0


def assert_close(a, e, tol, what=""):
    err = rel_err(a, e)
    assert err <= tol, f"{what}: rel err {err:.3e} > {tol:.1e}\n got {a}\n exp {e}"


def spec_from_leaves(kind: str, leaves: list):
    """gprx_amd KernelSpec from a golden's leaf list (tests/goldens.rs:160-230)."""
    import gprx_amd as g

    def leaf(lf):
        t = lf["type"]
        if t == "constant":
            return g.ConstantKernel(lf["constant"])
        if t == "rbf":
            return g.RbfKernel(lf["lengthscale"])
        if t == "periodic":
            return g.PeriodicKernel(lf["lengthscale"], lf["period"])
        if t == "product":
            ks = [leaf(c) for c in lf["leaves"]]
            out = ks[0]
            for k in ks[1:]:
                out = out * k
            return out
        if t == "sum":
            ks = [leaf(c) for c in lf["leaves"]]
            out = ks[0]
            for k in ks[1:]:
                out = out + k
            return out
        raise KeyError(t)

    ks = [leaf(lf) for lf in leaves]
    if kind == "leaf":
        return ks[0]
    out = ks[0]
    for k in ks[1:]:
        out = (out + k) if kind == "sum" else (out * k)
    return out


EXACT_COMPOSITE = ["rbf_sum_n2", "constant_rbf_product_n2", "constant_rbf_plus_rbf_n2", "periodic_n2"]
EXACT_RBF = ["rbf_n2", "rbf_n3"]


def golden_problem(name: str):
    """(spec, theta incl. log noise, x, n, d, y, xs, q, golden dict)."""
    import gprx_amd as g

    gd = load_golden(name)
    if name in EXACT_RBF or gd.get("kernel") == "rbf":
        if "leaves" in gd:
            spec = spec_from_leaves(gd["kernel"], gd["leaves"])
        else:
            ls = gd["lengthscales"][0] if "lengthscales" in gd else gd["lengthscale"]
            spec = g.RbfKernel(ls)
    elif gd.get("kernel") == "rbf_ard":
        spec = g.RbfArdKernel(gd["lengthscales"])
    else:
        spec = spec_from_leaves(gd["kernel"], gd["leaves"])
    theta = np.concatenate([spec.get_params(), [np.log(gd["noise_variance"])]])
    n, d = gd["n_rows"], gd["n_cols"]
    return spec, theta, np.asarray(gd["x"]), n, d, np.asarray(gd["y"]), np.asarray(gd["xs"]), gd["xs_n_rows"], gd


def has_gpu() -> bool:
    try:
        from gprx_amd import _lib

        return _lib.lib().gprx_hip_device_count() > 0
    except Exception:
        return False


@pytest.fixture(scope="session")
def gpu():
    if not has_gpu():
        pytest.skip("no GPU / library")
    return True


def duplicate_start_problem(n=300, d=3, seed=21):
    """Points 0 and 1 coincide, RBF lengthscale 0.05 on uniform points,
    noise 1e-10.  Rounded to f32, K + sigma^2 I holds 1.0 at (0,0), (0,1)
    and (1,1), so its second pivot is exactly 1 - 1*1 = 0 and an f32 factor
    fails without jitter; in f64 that pivot is 2e-10 and the factor holds.
    Returns (spec, noise, x column-major, n, d, y)."""
    import gprx_amd as g

    rng = np.random.default_rng(seed)
    pts = rng.random((n, d))
    pts[1] = pts[0]
    y = np.sin(6 * pts[:, 0]) + 0.1 * rng.standard_normal(n)
    return g.RbfKernel(0.05), 1e-10, pts.T.copy().ravel(), n, d, y
