"""gprx_hip_exact_predict_noalloc through the C ABI: the same predictions as
gprx_hip_exact_predict (FP64 to 1e-12, the mixed form to 1e-6, both
|a - e| / max(|e|, 1)) on two calls in a row, the second with fewer query
points; device memory unchanged across its calls; its refusals."""
import ctypes as C
import math

import numpy as np
import pytest

import gprx_amd as g
from gprx_amd import _lib as L

pytestmark = pytest.mark.gpu


def _problem(n=700, d=3, seed=8):
    rng = np.random.default_rng(seed)
    x = rng.random(n * d)
    y = np.sin(4 * x[:n]) + np.cos(3 * x[n:2 * n]) + 0.05 * rng.standard_normal(n)
    spec = g.ConstantKernel(1.1) * g.RbfArdKernel([0.4, 0.6, 0.8]) + g.RbfKernel(1.3)
    return spec, x, n, d, y


def _predict(fn, h, xs, q, d, kind, with_var=True):
    mean = np.full(q, np.nan)
    var = np.full(q, np.nan) if with_var else None
    err = L.GprxError()
    rc = fn(h, L.dptr(xs), q, d, kind, L.dptr(mean), L.dptr(var) if with_var else None, C.byref(err))
    return rc, mean, var, err


def _free_bytes():
    hip = C.CDLL("libamdhip64.so")
    free, total = C.c_size_t(), C.c_size_t()
    assert hip.hipMemGetInfo(C.byref(free), C.byref(total)) == 0
    return free.value


def _rel(a, e):
    return float(np.max(np.abs(a - e) / np.maximum(np.abs(e), 1.0)))


@pytest.mark.parametrize("prec,tol", [(g.Precision.FP64, 1e-12), (g.Precision.MIXED, 1e-6)])
def test_noalloc_predict_matches_predict(gpu, prec, tol):
    spec, x, n, d, y = _problem()
    m = g.Gpr(spec, g.GaussianLikelihood(0.02)).with_precision(prec).factor(x, n, d, y)
    lib = L.lib()
    rng = np.random.default_rng(3)
    for q in (300, 57):  # three chunks, then part of one
        xs = rng.random(q * d)
        for kind in (g.VarianceKind.OBSERVATION, g.VarianceKind.LATENT):
            print(f"[noalloc] precision {prec} q={q} kind {kind}", flush=True)
            rc0, m0, v0, e0 = _predict(lib.gprx_hip_exact_predict, m._h, xs, q, d, kind)
            rc1, m1, v1, e1 = _predict(lib.gprx_hip_exact_predict_noalloc, m._h, xs, q, d, kind)
            assert rc0 == 0 and rc1 == 0, (e0.message, e1.message)
            print(f"  mean err {_rel(m1, m0):.2e} variance err {_rel(v1, v0):.2e}", flush=True)
            assert _rel(m1, m0) <= tol and _rel(v1, v0) <= tol
        rc2, m2, _, _ = _predict(lib.gprx_hip_exact_predict_noalloc, m._h, xs, q, d, 0, with_var=False)
        assert rc2 == 0 and np.array_equal(m2, m1)
    m.close()


def test_noalloc_predict_allocates_nothing(gpu):
    spec, x, n, d, y = _problem(n=1000)
    m = g.Gpr(spec, g.GaussianLikelihood(0.02)).factor(x, n, d, y)
    xs = np.random.default_rng(4).random(500 * d)
    fn = L.lib().gprx_hip_exact_predict_noalloc
    assert _predict(fn, m._h, xs, 500, d, 0)[0] == 0
    before = _free_bytes()
    for q in (500, 130, 1):
        assert _predict(fn, m._h, xs, q, d, 0)[0] == 0
    assert _free_bytes() == before
    m.close()


def test_noalloc_predict_refusals(gpu):
    spec, x, n, d, y = _problem(n=200)
    fn = L.lib().gprx_hip_exact_predict_noalloc
    m = g.Gpr(spec, g.GaussianLikelihood(0.02)).prepare(x, n, d, y)  # not factored
    rc, *_ = _predict(fn, m._h, x[:5 * d], 5, d, 0)
    assert rc == L.INVALID_CONFIG
    m.close()
    big = 16384 + 1  # pads to 16512 rows, beyond the scratch
    xb = np.random.default_rng(5).random(big)
    mb = g.Gpr(g.RbfKernel(0.5), g.GaussianLikelihood(0.1)).prepare(xb, big, 1, np.sin(xb))
    rc, *_ = _predict(fn, mb._h, xb[:4], 4, 1, 0)
    assert rc == L.UNSUPPORTED
    mb.close()
