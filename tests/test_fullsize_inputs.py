"""The seeded generators still produce the inputs the full-size fixtures
(tests/golden/fullsize_*.json) were computed on (CPU only): first values of
every input and the parameter vectors, bit for bit."""
import json
import math

import numpy as np

from conftest import GOLDEN
from gprx_amd.configs import batched_problems, exact_ard, sgpr_vfe


def _fx(name):
    return json.loads((GOLDEN / name).read_text())


def test_exact_fixture_inputs():
    fx = _fx("fullsize_exact_n16384.json")
    wl = exact_ard(n=fx["n"], d=fx["d"], q=fx["q"], seed=2)
    assert wl.x[:8].tolist() == fx["x_head"] and wl.y[:8].tolist() == fx["y_head"]
    assert wl.xs[:8].tolist() == fx["xs_head"] and wl.theta().tolist() == fx["log_theta"]
    assert len(fx["grad"]) == 10 and len(fx["mean"]) == fx["q"]
    assert max(fx["oracle_vs_lapack"].values()) <= 1e-10


def test_sgpr_fixture_inputs():
    fx = _fx("fullsize_sgpr_n1048576.json")
    wl = sgpr_vfe(n=fx["n"], d=fx["d"], m=fx["m"], seed=3)
    assert wl.x[:8].tolist() == fx["x_head"] and wl.y[:8].tolist() == fx["y_head"]
    assert wl.z[:8].tolist() == fx["z_head"] and wl.theta().tolist() == fx["log_theta"]
    assert len(fx["grad"]) == 12 and fx["oracle_vs_lapack"]["value"] <= 1e-9
    assert fx["central_difference_err"] <= 1e-5


def test_batch_fixture_inputs():
    fx = _fx("fullsize_batch_subset.json")
    raw = batched_problems(k=256, seed=5)
    assert sorted({e["d"] for e in fx["problems"]}) == [2, 4, 8, 16]
    for e in fx["problems"]:
        n, d, x, y, ls, noise = raw[e["index"]]
        assert (n, d) == (e["n"], e["d"]) and x[:4].tolist() == e["x_head"] and y[:4].tolist() == e["y_head"]
        assert np.concatenate([np.log(ls), [math.log(noise)]]).tolist() == e["log_theta"]
