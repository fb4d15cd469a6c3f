"""Writes the full-size FP64 fixtures of the seeded workloads from the CPU oracle.

    python tests/golden/make_fullsize.py exact   # fullsize_exact_n16384.json
    python tests/golden/make_fullsize.py sgpr    # fullsize_sgpr_n1048576.json
    python tests/golden/make_fullsize.py batch   # fullsize_batch_subset.json

Each fixture holds the oracle's results (oracle/gprx_oracle.c, f64) and, as a
check of the oracle at that size, the same quantities from an independent
LAPACK/BLAS formulation (scipy/numpy: Cholesky, triangular solves and
explicit inverse), with their largest relative difference |a - e| / max(|e|, 1).
The script refuses to write a fixture whose two formulations differ by more
than 1e-10 (exact; 1e-9 on the gradients of the batched problems, whose
condition numbers reach 1e7) or 1e-9 (VFE).  The first values of each generated input
are stored too, so a test can tell a changed generator from a changed result.

Sizes and workloads:
  exact  -- Constant(1) x RbfArd, n = 16384, d = 8, seed 2, q = 256 test
            points: NLML, the 10 gradient components, mean, latent and
            observation variance.  Oracle: about 5 minutes on 8 threads.
  sgpr   -- Constant x RbfArd + Periodic, n = 2^20, d = 8, m = 1024, seed 3:
            the VFE bound and its 12 gradient components; the LAPACK form
            checks the bound, and central differences of it check 3
            gradient components.  Oracle: about 6 minutes on 8 threads.
  batch  -- every 16th of the 256 problems of the batched workload (seed 5,
            RbfArd, n in [256, 4096], d in {2, 4, 8, 16}): NLML and gradient.
"""
from __future__ import annotations

import json
import math
import sys
import time
from pathlib import Path

import numpy as np
import scipy.linalg as sl

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

import oracle  # noqa: E402
from gprx_amd.configs import batched_problems, exact_ard, sgpr_vfe  # noqa: E402

OUT = Path(__file__).resolve().parent


def rel(a, e) -> float:
    a, e = np.asarray(a, float), np.asarray(e, float)
    return float(np.max(np.abs(a - e) / np.maximum(np.abs(e), 1.0)))


def log(msg: str) -> None:
    print(f"[{time.strftime('%H:%M:%S')}] {msg}", flush=True)


def lapack_exact(ops, th, x, n, d, y, xs, q):
    """NLML, gradient and predictions with LAPACK's Cholesky and inverse; the
    kernel values and the <W, dK> contraction are the oracle's (pinned by the
    reference's goldens at small n)."""
    noise = math.exp(th[-1])
    K = oracle.cross(ops, th, x, n, x, n, d)
    K[np.diag_indices(n)] += noise
    c = sl.cho_factor(K, lower=True, overwrite_a=True, check_finite=False)
    alpha = sl.cho_solve(c, y, check_finite=False)
    logdet = float(np.sum(np.log(np.diag(c[0]))))
    nlml = 0.5 * float(y @ alpha) + logdet + 0.5 * n * math.log(2 * math.pi)
    Ai = sl.cho_solve(c, np.eye(n), check_finite=False)
    W = np.outer(alpha, alpha) - Ai
    del Ai
    grad = np.empty(th.size)
    grad[:-1] = -0.5 * oracle.contract(ops, th, x, n, x, n, d, W, sym=True)
    grad[-1] = -0.5 * noise * float(np.trace(W))
    del W
    Ks = oracle.cross(ops, th, x, n, xs, q, d)
    mean = Ks.T @ alpha
    V = sl.solve_triangular(c[0], Ks, lower=True, check_finite=False)
    kss = np.array([oracle.cross(ops, th, xs[j::q], 1, xs[j::q], 1, d)[0, 0] for j in range(q)])
    lat = np.maximum(kss - np.sum(V * V, axis=0), 0.0)
    return nlml, grad, mean, lat, lat + noise


def make_exact() -> None:
    wl = exact_ard(n=16384, d=8, q=256, seed=2)
    ops, th = wl.kernel().program(), wl.theta()
    log(f"exact n={wl.n} d={wl.d} q={wl.q}: oracle value and gradient")
    t = time.time()
    nlml, grad, _, _ = oracle.exact_value_grad(ops, th, wl.x, wl.n, wl.d, wl.y, blocked=True)
    t_vg = time.time() - t
    log(f"  oracle NLML {nlml!r} in {t_vg:.1f} s; predict")
    mean, lat, obs = oracle.exact_predict(ops, th, wl.x, wl.n, wl.d, wl.y, wl.xs, wl.q)
    log("  LAPACK formulation")
    lv, lg, lm, ll, lo = lapack_exact(ops, th, wl.x, wl.n, wl.d, wl.y, wl.xs, wl.q)
    diff = {"nlml": rel(nlml, lv), "grad": rel(grad, lg), "mean": rel(mean, lm),
            "latent_variance": rel(lat, ll), "observation_variance": rel(obs, lo)}
    log(f"  oracle vs LAPACK: {diff}")
    assert max(diff.values()) <= 1e-10, diff
    doc = {
        "workload": "exact GPR, Constant(1) x RbfArd, gprx_amd.configs.exact_ard(n=16384, d=8, q=256, seed=2)",
        "n": wl.n, "d": wl.d, "q": wl.q, "log_theta": th.tolist(),
        "x_head": wl.x[:8].tolist(), "y_head": wl.y[:8].tolist(), "xs_head": wl.xs[:8].tolist(),
        "nlml": nlml, "grad": grad.tolist(), "mean": mean.tolist(), "latent_variance": lat.tolist(),
        "observation_variance": obs.tolist(),
        "oracle_vs_lapack": diff, "oracle_seconds": t_vg, "oracle_threads": oracle.threads(),
    }
    (OUT / "fullsize_exact_n16384.json").write_text(json.dumps(doc, indent=1) + "\n")
    log("wrote fullsize_exact_n16384.json")


def lapack_vfe_bound(ops, th, x, n, d, y, z, m) -> float:
    """The VFE bound (src/sgpr/factor/vfe.rs:191-369) with LAPACK/BLAS dense
    algebra and the adaptive K_mm jitter (1e-8, x10, 5 tries, <= 1e-3)."""
    noise = math.exp(th[-1])
    Kmm = oracle.cross(ops, th, z, m, z, m, d)
    jit, nxt = 0.0, 1e-8
    for _ in range(6):
        try:
            L = np.linalg.cholesky(Kmm + jit * np.eye(m))
            break
        except np.linalg.LinAlgError:
            jit, nxt = nxt, nxt * 10
    else:
        raise RuntimeError("K_mm factor failed")
    A = oracle.cross(ops, th, z, m, x, n, d)
    A = sl.solve_triangular(L, A, lower=True, overwrite_b=True, check_finite=False)
    B = A @ A.T
    B[np.diag_indices(m)] += noise
    LB = np.linalg.cholesky(B)
    Ay = A @ y
    w = sl.cho_solve((LB, True), Ay, check_finite=False)
    # k(x, x) of a stationary kernel (RBF-ARD, periodic of a distance) does
    # not depend on x: one entry times n.
    kdiag_sum = float(oracle.cross(ops, th, x[::n], 1, x[::n], 1, d)[0, 0]) * n
    a_f2 = float(np.sum(A * A))
    logdet = (n - m) * math.log(noise) + 2.0 * float(np.sum(np.log(np.diag(LB))))
    quad = (float(y @ y) - float(Ay @ w)) / noise
    trace = (kdiag_sum - a_f2) / (2.0 * noise)
    return 0.5 * (n * math.log(2 * math.pi) + logdet + quad) + trace


def make_sgpr() -> None:
    wl = sgpr_vfe(n=1 << 20, d=8, m=1024, seed=3)
    ops, th = wl.kernel().program(), wl.theta()
    log(f"sgpr n={wl.n} m={wl.m} d={wl.d}: oracle bound and gradient")
    t = time.time()
    val, grad = oracle.vfe_value_grad(ops, th, wl.x, wl.n, wl.d, wl.y, wl.z, wl.m)
    t_vg = time.time() - t
    log(f"  oracle bound {val!r} in {t_vg:.1f} s; LAPACK bound")
    lv = lapack_vfe_bound(ops, th, wl.x, wl.n, wl.d, wl.y, wl.z, wl.m)
    diff = {"value": rel(val, lv)}
    log(f"  oracle vs LAPACK bound: {diff}")
    assert diff["value"] <= 1e-9, diff
    h = 1e-4
    fd = {}
    for k in (0, 1, th.size - 1):
        e = np.zeros_like(th)
        e[k] = h
        fp = lapack_vfe_bound(ops, th + e, wl.x, wl.n, wl.d, wl.y, wl.z, wl.m)
        fm = lapack_vfe_bound(ops, th - e, wl.x, wl.n, wl.d, wl.y, wl.z, wl.m)
        fd[str(k)] = (fp - fm) / (2 * h)
        log(f"  d/dtheta[{k}] oracle {grad[k]!r} central difference {fd[str(k)]!r}")
    fd_err = max(abs(fd[str(k)] - grad[k]) / max(abs(grad[k]), 1.0) for k in (0, 1, th.size - 1))
    assert fd_err <= 1e-5, fd_err
    doc = {
        "workload": "Sgpr VFE, Constant x RbfArd + Periodic, gprx_amd.configs.sgpr_vfe(n=2**20, d=8, m=1024, seed=3)",
        "n": wl.n, "d": wl.d, "m": wl.m, "log_theta": th.tolist(),
        "x_head": wl.x[:8].tolist(), "y_head": wl.y[:8].tolist(), "z_head": wl.z[:8].tolist(),
        "value": val, "grad": grad.tolist(),
        "oracle_vs_lapack": diff, "central_differences": fd, "central_difference_h": h,
        "central_difference_err": fd_err, "oracle_seconds": t_vg, "oracle_threads": oracle.threads(),
    }
    (OUT / "fullsize_sgpr_n1048576.json").write_text(json.dumps(doc, indent=1) + "\n")
    log("wrote fullsize_sgpr_n1048576.json")


def make_batch() -> None:
    raw = batched_problems(k=256, seed=5)
    import gprx_amd as g

    out = []
    worst = worst_g = 0.0
    for i in range(0, 256, 16):
        n, d, x, y, ls, noise = raw[i]
        spec = g.RbfArdKernel(list(ls))
        th = np.concatenate([np.log(ls), [math.log(noise)]])
        ops = spec.program()
        v, gr, _, _ = oracle.exact_value_grad(ops, th, x, n, d, y, blocked=True)
        lv, lg, *_ = lapack_exact(ops, th, x, n, d, y, x[::n], 1)
        dv, dg = rel(v, lv), rel(gr, lg)
        worst = max(worst, dv)
        worst_g = max(worst_g, dg)
        log(f"  problem {i}: n={n} d={d} nlml {v!r} oracle vs LAPACK: value {dv:.2e} grad {dg:.2e}")
        out.append({"index": i, "n": n, "d": d, "log_theta": th.tolist(), "x_head": x[:4].tolist(),
                    "y_head": y[:4].tolist(), "nlml": v, "grad": gr.tolist(),
                    "oracle_vs_lapack": {"nlml": dv, "grad": dg}})
    # The noise variances reach 1e-3 with d = 2: kappa up to about 1e7, so the
    # gradient agrees to kappa * eps, not to 1e-10.
    assert worst <= 1e-10 and worst_g <= 1e-9, (worst, worst_g)
    doc = {"workload": "batched exact GPR, RbfArd, gprx_amd.configs.batched_problems(k=256, seed=5); "
                       "every 16th problem",
           "problems": out, "oracle_vs_lapack_max": {"nlml": worst, "grad": worst_g}}
    (OUT / "fullsize_batch_subset.json").write_text(json.dumps(doc, indent=1) + "\n")
    log("wrote fullsize_batch_subset.json")


if __name__ == "__main__":
    oracle.build()
    which = sys.argv[1:] or ["exact", "sgpr", "batch"]
    for w in which:
        {"exact": make_exact, "sgpr": make_sgpr, "batch": make_batch}[w]()
