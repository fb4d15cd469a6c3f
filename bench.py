#!/usr/bin/env python3
"""NLML+grad evaluations/s, exact GPR n=16384 d=8 ARD-RBF, FP32-mixed vs FP64, one MI355X.

A step is one NLML + gradient evaluation (all 10 log-parameters) at fixed
parameters, with the training data resident in HBM: the reference's own perf
clock (FittedGpr::value_and_gradient_into after Gpr<Fixed>::factor).
`value` is the mixed form's evaluations/s; the FP64 form (the correctness
oracle form on the GPU) is timed in the same run and reported beside it with
the mixed form's errors against it.

Every library call returns after a stream synchronisation, so the host clock
around K calls brackets exactly K evaluations.

    python bench.py [--steps K] [--warmup W] [--n N] [--gpus 1]
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

PEAK_FP64 = 78.6  # TFLOP/s, MI355X dense FP64 matrix
PEAK_FP32 = 157.3  # TFLOP/s, MI355X dense FP32 matrix
PEAK_HBM = 8000.0  # GB/s


def exact_flops(n: int, d: int) -> float:
    """Algorithmic FLOP of one exact NLML+grad eval: factor n^3/3, inverse
    2n^3/3, alpha 2n^2, fill 3d n(n+1)/2, gradient fold 2 n^2 d."""
    return n ** 3 / 3 + 2 * n ** 3 / 3 + 2 * n ** 2 + 3 * d * n * (n + 1) / 2 + 2 * n * n * d


def inverse_product_flops(n: int) -> float:
    """X^T X over the lower triangle: sum_i (i+1) 2 (n-i) = n(n+1)(n+2)/3."""
    return n * (n + 1) * (n + 2) / 3


def rel(a, e):
    a = np.asarray(a, float)
    e = np.asarray(e, float)
    return float(np.max(np.abs(a - e) / np.maximum(np.abs(e), 1e-300)))


def time_form(gx, wl, precision, steps, warmup):
    model = gx.Gpr(wl.kernel(), gx.GaussianLikelihood(wl.noise)).with_precision(precision).prepare(
        wl.x, wl.n, wl.d, wl.y)
    th = wl.theta()
    grad = np.empty(th.size)
    for _ in range(warmup):
        model.value_and_gradient_into(th, grad)
    phases = np.zeros(5)
    evs = []
    t0 = time.perf_counter()
    for _ in range(steps):
        val = model.value_and_gradient_into(th, grad)
        phases += np.asarray(model.last_phase_ms())
        evs.append(model.last_eval_ms())
    wall = time.perf_counter() - t0
    return model, val, grad.copy(), wall, phases / steps, float(np.mean(evs))


def cpu_baseline(wl, cpu_n):
    """The oracle (a restatement of gprx's CPU algorithm: Cholesky, solve(L, I),
    the W contraction) on the first cpu_n points, all host threads, one
    evaluation; scaled to n by the n^3 cost."""
    import oracle

    n0 = min(cpu_n, wl.n)
    X = wl.x.reshape(wl.d, wl.n)[:, :n0].reshape(-1)
    y = wl.y[:n0]
    ops = wl.kernel().program()
    th = wl.theta()
    t0 = time.perf_counter()
    oracle.exact_value_grad(ops, th, X, n0, wl.d, y, blocked=True)
    t = time.perf_counter() - t0
    scale = (n0 / wl.n) ** 3
    return {
        "value": scale / t,
        "unit": "evals/s",
        "cores": oracle.threads(),
        "kind": "port",
        "sample": f"one NLML+grad eval of the oracle at n={n0} d={wl.d} ({t:.2f} s), "
                  f"scaled to n={wl.n} by (n/{n0})^3",
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=16384)
    ap.add_argument("--cpu-n", type=int, default=4096)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-predict", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: this path runs on one GPU; exact GPR at n=16384 does not shard "
              "across GPUs here (see DESIGN.md, sharding).", file=sys.stderr)
        sys.exit(2)

    import gprx_amd as gx
    from gprx_amd.configs import exact_ard

    wl = exact_ard(n=args.n)
    print(f"[bench] exact ARD n={wl.n} d={wl.d}: timing FP64", file=sys.stderr, flush=True)
    m64, v64, g64, wall64, ph64, ev64 = time_form(gx, wl, gx.Precision.FP64, args.steps, args.warmup)
    print(f"[bench] FP64 {wall64 / args.steps * 1e3:.1f} ms/eval; timing MIXED", file=sys.stderr,
          flush=True)
    m32, v32, g32, wall32, ph32, ev32 = time_form(gx, wl, gx.Precision.MIXED, args.steps, args.warmup)
    print(f"[bench] MIXED {wall32 / args.steps * 1e3:.1f} ms/eval", file=sys.stderr, flush=True)

    evals64 = args.steps / wall64
    evals32 = args.steps / wall32
    F = exact_flops(wl.n, wl.d)
    lauum = inverse_product_flops(wl.n)
    # Dominant kernel: the X^T X tile-GEMM launch (one launch per evaluation).
    ach32 = lauum / (ph32[3] * 1e-3) / 1e12
    ach64 = lauum / (ph64[3] * 1e-3) / 1e12

    extra = {}
    if not args.no_predict:
        t0 = time.perf_counter()
        p64 = m64.predict(wl.xs, wl.q, wl.d)
        t64 = time.perf_counter() - t0
        t0 = time.perf_counter()
        p32 = m32.predict(wl.xs, wl.q, wl.d)
        t32 = time.perf_counter() - t0
        extra["predict_ms"] = {"fp64": t64 * 1e3, "mixed": t32 * 1e3, "q": wl.q}
        extra["mixed_predict_rel_err"] = {"mean": rel(p32.mean, p64.mean), "variance": rel(p32.variance, p64.variance)}

    out = {
        "metric": "NLML+grad evals/s, exact GPR n=16384 d=8 ARD-RBF: FP32-mixed vs FP64 on GPU",
        "value": evals32,
        "unit": "evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall32 / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32-mixed (f64 panels, log-det, refinement)",
        "data": "synthetic (seeded SeededRng, seed 2)",
        "config": {"workload": "exact GPR NLML+grad, Constant x RbfArd", "n": wl.n, "d": wl.d,
                   "params": int(wl.theta().size), "parallelism": "single GPU"},
        "fp64_evals_per_s": evals64,
        "fp64_ms_per_step": wall64 / args.steps * 1e3,
        "mixed_vs_fp64": evals32 / evals64,
        "mixed_rel_err_vs_fp64": {"nlml": rel(v32, v64), "grad_max": rel(g32, g64),
                                  "grad": [float(v) for v in np.abs(g32 - g64) / np.maximum(np.abs(g64), 1e-300)]},
        "mixed_refine_steps": m32.last_refine_steps(),
        "mixed_fallback": m32.last_fallback(),
        "phase_ms": {"names": ["fill", "factor+inverse", "alpha", "XtX", "contract"],
                     "mixed": [round(v, 3) for v in ph32], "fp64": [round(v, 3) for v in ph64]},
        "eval_tflops": {"mixed": F / (wall32 / args.steps) / 1e12, "fp64": F / (wall64 / args.steps) / 1e12,
                        "algorithmic_flop_per_eval": F},
        "roofline": {"bound": "mfma", "kernel": "k_gemm<float,T,N> (X^T X, one launch per eval)",
                     "achieved": ach32, "peak": PEAK_FP32, "unit": "TFLOP/s", "frac": ach32 / PEAK_FP32,
                     "traffic": None},
        "roofline_fp64": {"bound": "mfma", "kernel": "k_gemm<double,T,N> (X^T X)", "achieved": ach64,
                          "peak": PEAK_FP64, "unit": "TFLOP/s", "frac": ach64 / PEAK_FP64},
        **extra,
    }
    if not args.no_cpu_baseline:
        print(f"[bench] CPU baseline (oracle, n={args.cpu_n})", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline(wl, args.cpu_n)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
