#!/usr/bin/env python3
"""NLML+grad evaluations/s, exact GPR n=16384 d=8 ARD-RBF, FP32-mixed vs FP64, one MI355X.

A step is one NLML + gradient evaluation (all 10 log-parameters) at fixed
parameters, with the training data resident in HBM: the reference's own perf
clock (FittedGpr::value_and_gradient_into after Gpr<Fixed>::factor).
`value` is the FP64 form's evaluations/s: the form that meets its accuracy
bar (1e-10 against the CPU f64 path).  The mixed form (f32 MFMA factor and
inverse, f64 refinement of alpha) is timed in the same run and reported in
`mixed` with its errors against FP64; it does not meet its 1e-8 / 1e-6 bar.
The 256-problem batch of the batched workload is timed beside it, batched
against the looped single-problem path (`batch256`).

Every library call returns after a stream synchronisation, so the host clock
around K calls brackets exactly K evaluations.

    python bench.py [--steps K] [--warmup W] [--n N] [--gpus 1]

Other workloads of the same path, each printing one JSON line of its own:
    --workload sgpr    Sgpr (VFE) bound + gradient, n=2^20 d=8 m=1024
    --workload svgp    Svgp Adam steps, n=2^22 d=16 m=2048 b=4096
    --workload batch   256 exact problems of uneven n, batched vs looped
    --workload copy    a plain 1 GiB device copy (calibrates the HBM byte
                       counters of rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE)
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time
from pathlib import Path

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

PEAK_FP64 = 78.6  # TFLOP/s, MI355X dense FP64 matrix
PEAK_FP32 = 157.3  # TFLOP/s, MI355X dense FP32 matrix
PEAK_HBM = 8000.0  # GB/s

# HBM-side bytes per X^T X launch at n = 16384: rocprofv3 --pmc FETCH_SIZE and
# WRITE_SIZE on that launch, each scaled by what the counter reads for a 1 GiB
# device copy (x2.0 and x1.0); the run is summarised under profiles/.
XTX_TRAFFIC_N16384 = {"mixed": 3.167e+10, "fp64": 6.639e+10}
# HBM-side bytes of one whole evaluation at n = 16384: FETCH_SIZE + WRITE_SIZE
# summed over the evaluation's dispatches, scaled as above (the model's
# one-time buffer clears left out).
EVAL_TRAFFIC_N16384 = {"fp64": 1.557e11}


def exact_flops(n: int, d: int) -> float:
    """Algorithmic FLOP of one exact NLML+grad eval: factor n^3/3, inverse
    2n^3/3, alpha 2n^2, fill 3d n(n+1)/2, gradient fold 2 n^2 d."""
    return n ** 3 / 3 + 2 * n ** 3 / 3 + 2 * n ** 2 + 3 * d * n * (n + 1) / 2 + 2 * n * n * d


def inverse_product_flops(n: int) -> float:
    """X^T X over the lower triangle: sum_i (i+1) 2 (n-i) = n(n+1)(n+2)/3."""
    return n * (n + 1) * (n + 2) / 3


def sgpr_flops(n: int, m: int) -> float:
    """Algorithmic FLOP of one VFE bound + gradient: the three m x m x n
    contractions a minimal formulation needs (K_mn K_mn^T or A A^T, the
    whitening product and the weight product), 3 m^2 n; the m^3 terms are
    below 0.5% at n = 1024 m."""
    return 3.0 * m * m * n


def svgp_flops(b: int, m: int) -> float:
    """Algorithmic FLOP of one SVGP step: 6 m^2 b for the b-long products
    (A, U = Lq^T A, A U^T) and 7 m^3 / 3 for K_mm's factor and the m x m
    algebra."""
    return 6.0 * m * m * b + 7.0 * m ** 3 / 3


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


def cpu_baseline(wl, warmup, reps):
    """The oracle (a restatement of gprx's CPU algorithm: Cholesky,
    solve(L, I), the W contraction) on the whole workload, all host threads,
    on the reference's clock (compare/perf/gprx/src/timing.rs: discard
    `warmup` evaluations, then the median of `reps`).  One evaluation at
    n = 16384 takes about a minute on 16 host threads, so the default is
    PERF_WARMUP = 1 and PERF_REPS = 1, settings of that clock; each finished
    phase prints a line."""
    import oracle

    ops = wl.kernel().program()
    th = wl.theta()
    oracle.set_verbose(True)
    samples = []
    for i in range(warmup + reps):
        print(f"[bench] CPU baseline: oracle evaluation {i + 1} of {warmup + reps} at n={wl.n}",
              file=sys.stderr, flush=True)
        t0 = time.perf_counter()
        oracle.exact_value_grad(ops, th, wl.x, wl.n, wl.d, wl.y, blocked=True)
        if i >= warmup:
            samples.append(time.perf_counter() - t0)
    oracle.set_verbose(False)
    t = float(np.median(samples))
    return {
        "value": 1.0 / t,
        "unit": "evals/s",
        "cores": oracle.threads(),
        "kind": "port",
        "sample": f"the whole workload (n={wl.n} d={wl.d}): median of {reps} oracle NLML+grad "
                  f"evaluation(s) after {warmup} discarded, {t:.1f} s each; the oracle restates "
                  f"gprx's CPU algorithm (Cholesky, solve(L, I), W contraction) in C with OpenMP",
    }


def bench_exact(args, gx):
    from gprx_amd.configs import exact_ard

    wl = exact_ard(n=args.n)
    print(f"[bench] exact ARD n={wl.n} d={wl.d}: timing FP64", file=sys.stderr, flush=True)
    m64, v64, g64, wall64, ph64, ev64 = time_form(gx, wl, gx.Precision.FP64, args.steps, args.warmup)
    print(f"[bench] FP64 {wall64 / args.steps * 1e3:.1f} ms/eval; timing MIXED", file=sys.stderr,
          flush=True)
    if "mixed" not in args.forms.split(","):
        # profiling runs of the FP64 evaluation alone
        m64.close()
        return {"metric": "NLML+grad evals/s, exact GPR FP64 only", "value": args.steps / wall64,
                "unit": "evals/s", "ms_per_step": wall64 / args.steps * 1e3,
                "phase_ms": [round(v, 3) for v in ph64], "eval_ms_hip_events": ev64}
    m32, v32, g32, wall32, ph32, ev32 = time_form(gx, wl, gx.Precision.MIXED, args.steps, args.warmup)
    print(f"[bench] MIXED {wall32 / args.steps * 1e3:.1f} ms/eval", file=sys.stderr, flush=True)

    evals64 = args.steps / wall64
    evals32 = args.steps / wall32
    F = exact_flops(wl.n, wl.d)
    lauum = inverse_product_flops(wl.n)
    # The X^T X tile-GEMM launch (one per evaluation), timed with HIP events
    # on the model's stream.
    ach32 = lauum / (ph32[3] * 1e-3) / 1e12
    ach64 = lauum / (ph64[3] * 1e-3) / 1e12
    # The whole evaluation, from HIP events on the model's stream around it.
    eval64 = F / (ev64 * 1e-3) / 1e12
    eval32 = F / (ev32 * 1e-3) / 1e12

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
    e_nlml, e_grad = rel(v32, v64), rel(g32, g64)
    mixed = {
        "evals_per_s": evals32, "ms_per_step": wall32 / args.steps * 1e3,
        "vs_fp64": evals32 / evals64,
        "rel_err_vs_fp64": {"nlml": e_nlml, "grad_max": e_grad,
                            "grad": [float(v) for v in np.abs(g32 - g64) / np.maximum(np.abs(g64), 1e-300)]},
        "meets_target": {"nlml_le_1e-8": e_nlml <= 1e-8, "grad_le_1e-6": e_grad <= 1e-6,
                         "speedup_ge_3": evals32 / evals64 >= 3.0},
        "refine_steps": m32.last_refine_steps(), "fallback": m32.last_fallback(),
        "roofline": {"bound": "mfma", "kernel": "whole mixed evaluation", "achieved": eval32,
                     "peak": PEAK_FP32, "unit": "TFLOP/s", "frac": eval32 / PEAK_FP32},
    }
    m64.close()
    m32.close()
    out = {
        "metric": "NLML+grad evals/s, exact GPR n=16384 d=8 ARD-RBF: FP32-mixed vs FP64 on GPU",
        "value": evals64,
        "unit": "evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall64 / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded SeededRng, seed 2)",
        "config": {"workload": "exact GPR NLML+grad, Constant x RbfArd", "n": wl.n, "d": wl.d,
                   "params": int(wl.theta().size), "parallelism": "single GPU"},
        "value_form": "FP64 (meets 1e-10 against the CPU f64 oracle); the mixed form is in `mixed`",
        "mixed": mixed,
        "phase_ms": {"names": ["fill", "factor+inverse", "alpha", "XtX", "contract"],
                     "fp64": [round(v, 3) for v in ph64], "mixed": [round(v, 3) for v in ph32]},
        "algorithmic_flop_per_eval": F,
        "roofline": {"bound": "mfma",
                     "kernel": "whole FP64 evaluation, HIP events on its stream (by rocprof total time "
                               "its dominant kernel is k_gemm<double,N,T,128>, the factor's "
                               "L21 and trailing-update products, on three streams)",
                     "achieved": eval64, "peak": PEAK_FP64, "unit": "TFLOP/s", "frac": eval64 / PEAK_FP64,
                     "traffic": EVAL_TRAFFIC_N16384.get("fp64") if wl.n == 16384 else None},
        "roofline_xtx": {"bound": "mfma", "kernel": "k_gemm<double,T,N,128> (X^T X, one launch per eval)",
                         "achieved": ach64, "peak": PEAK_FP64, "unit": "TFLOP/s", "frac": ach64 / PEAK_FP64,
                         "traffic": XTX_TRAFFIC_N16384["fp64"] if wl.n == 16384 else None},
        "roofline_xtx_mixed": {"bound": "mfma", "kernel": "k_gemm<float,T,N,128> (X^T X)",
                               "achieved": ach32, "peak": PEAK_FP32, "unit": "TFLOP/s",
                               "frac": ach32 / PEAK_FP32},
        **extra,
    }
    if not args.no_batch:
        print("[bench] 256-problem batch (FP64): batched and looped", file=sys.stderr, flush=True)
        forms, Fb = time_batch(gx, ("fp64",), steps=2, warmup=1)
        fb = forms["fp64"]
        out["batch256"] = {"workload": "256 exact ARD-RBF problems, n in [256, 4096], d in {2,4,8,16} "
                                       "(seed 5), FP64",
                           "batched_ms": fb["batched_ms"], "looped_ms": fb["looped_ms"],
                           "batched_vs_looped": fb["speedup"],
                           "batched_vs_looped_err": fb["batched_vs_looped_err"],
                           "eval_tflops": Fb / (fb["batched_ms"] * 1e-3) / 1e12}
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(wl, args.cpu_warmup, args.cpu_reps)
    return out


def _timed(fn, steps, warmup):
    for _ in range(warmup):
        fn()
    t0 = time.perf_counter()
    for _ in range(steps):
        r = fn()
    return r, time.perf_counter() - t0


def bench_sgpr(args, gx):
    from gprx_amd.configs import sgpr_vfe

    wl = sgpr_vfe()
    th = wl.theta()
    res = {}
    for name, prec in (("fp64", gx.Precision.FP64), ("mixed", gx.Precision.MIXED)):
        print(f"[bench] sgpr n={wl.n} m={wl.m}: {name}", file=sys.stderr, flush=True)
        f = gx.Sgpr(wl.kernel(), gx.GaussianLikelihood(wl.noise)).with_precision(prec).prepare(
            wl.x, wl.n, wl.d, wl.y, wl.z, wl.m)
        grad = np.empty(th.size)
        val, wall = _timed(lambda: f.value_and_gradient_into(th, grad), args.steps, args.warmup)
        res[name] = (val, grad.copy(), wall)
        f.close()
    F = sgpr_flops(wl.n, wl.m)
    v64, g64, w64 = res["fp64"]
    v32, g32, w32 = res["mixed"]
    ach = F / (w32 / args.steps) / 1e12
    ach64 = F / (w64 / args.steps) / 1e12
    return {
        "metric": "VFE bound+grad evals/s, Sgpr n=2^20 d=8 m=1024 Constant x RbfArd + Periodic",
        "value": args.steps / w32, "unit": "evals/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": w32 / args.steps * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32-mixed (f64 m x m algebra)",
        "data": "synthetic (seeded SeededRng, seed 3)",
        "config": {"workload": "Sgpr VFE bound+grad", "n": wl.n, "d": wl.d, "m": wl.m,
                   "params": int(th.size), "parallelism": "single GPU"},
        "fp64_evals_per_s": args.steps / w64, "mixed_vs_fp64": w64 / w32,
        "mixed_rel_err_vs_fp64": {"bound": rel(v32, v64), "grad_max": rel(g32, g64)},
        "roofline": {"bound": "mfma", "kernel": "whole evaluation (algorithmic 3 m^2 n)",
                     "achieved": ach, "peak": PEAK_FP32, "unit": "TFLOP/s", "frac": ach / PEAK_FP32,
                     "traffic": None},
        "roofline_fp64": {"bound": "mfma", "kernel": "whole evaluation", "achieved": ach64,
                          "peak": PEAK_FP64, "unit": "TFLOP/s", "frac": ach64 / PEAK_FP64},
    }


def bench_svgp(args, gx):
    from gprx_amd.configs import svgp_minibatch

    wl = svgp_minibatch()
    th = wl.theta()
    b, lr = 4096, 0.01
    steps = args.steps
    res = {}
    for name, prec in (("fp64", gx.Precision.FP64), ("mixed", gx.Precision.MIXED)):
        print(f"[bench] svgp n={wl.n} m={wl.m} b={b}: {name}, {steps} steps", file=sys.stderr, flush=True)
        f = gx.Svgp(wl.kernel(), gx.GaussianLikelihood(wl.noise)).with_precision(prec).prepare(
            wl.x, wl.n, wl.d, wl.y, wl.z, wl.m)
        p0 = f.params.copy()
        if args.warmup:
            f.adam(p0, steps=args.warmup, batch_size=b, lr=lr, seed=0)
        t0 = time.perf_counter()
        p, vals = f.adam(p0, steps=steps, batch_size=b, lr=lr, seed=0)
        res[name] = (vals, time.perf_counter() - t0)
        f.close()
    (v64, w64), (v32, w32) = res["fp64"], res["mixed"]
    F = svgp_flops(b, wl.m)
    ach = F / (w32 / steps) / 1e12
    ach64 = F / (w64 / steps) / 1e12
    traj = np.abs(v32 - v64) / np.maximum(np.abs(v64), 1e-300)
    return {
        "metric": "Adam steps/s, Svgp n=2^22 d=16 m=2048 b=4096 ARD-RBF",
        "value": steps / w32, "unit": "steps/s", "n_gpus": 1, "steps": steps,
        "warmup": args.warmup, "ms_per_step": w32 / steps * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32-mixed (f64 m x m algebra)",
        "data": "synthetic (seeded SeededRng, seed 4)",
        "config": {"workload": "Svgp minibatch Adam, lr 0.01", "n": wl.n, "d": wl.d, "m": wl.m,
                   "batch": b, "params": int(len(gx.prior_params(th, wl.m))),
                   "parallelism": "single GPU"},
        "fp64_steps_per_s": steps / w64, "mixed_vs_fp64": w64 / w32,
        "neg_elbo": {"fp64_first": float(v64[0]), "fp64_last": float(v64[-1]),
                     "mixed_last": float(v32[-1])},
        "mixed_elbo_traj_rel_err": {"max": float(traj.max()), "last": float(traj[-1])},
        "roofline": {"bound": "mfma", "kernel": "whole step (algorithmic 6 m^2 b + 7 m^3 / 3)",
                     "achieved": ach, "peak": PEAK_FP32, "unit": "TFLOP/s", "frac": ach / PEAK_FP32,
                     "traffic": None},
        "roofline_fp64": {"bound": "mfma", "kernel": "whole step", "achieved": ach64,
                          "peak": PEAK_FP64, "unit": "TFLOP/s", "frac": ach64 / PEAK_FP64},
    }


def time_batch(gx, forms, steps, warmup):
    """The 256-problem batched workload in each precision of `forms`: the
    batched evaluation and the looped single-problem path, each timed over
    `steps` calls after `warmup`.  Returns ({form: timings and errors},
    algorithmic FLOP of one batch)."""
    from gprx_amd.configs import batched_problems

    raw = batched_problems(k=256)
    probs = [(gx.RbfArdKernel(list(ls)), gx.GaussianLikelihood(noise), x, n, d, y)
             for (n, d, x, y, ls, noise) in raw]
    thetas = [np.concatenate([np.log(ls), [np.log(noise)]]) for (n, d, x, y, ls, noise) in raw]
    F = sum(exact_flops(n, d) for (n, d, *_r) in raw)
    out_forms = {}
    ref = None  # FP64 batched results: the reference the mixed form is held to

    def err1(a, e):
        a, e = np.asarray(a, float), np.asarray(e, float)
        return float(np.max(np.abs(a - e) / np.maximum(np.abs(e), 1.0)))

    precs = {"fp64": gx.Precision.FP64, "mixed": gx.Precision.MIXED}
    for name in forms:
        prec = precs[name]
        print(f"[bench] batch of {len(probs)}: {name} batched", file=sys.stderr, flush=True)
        bt = gx.GprBatch(probs, precision=prec)
        (nl, gr, er), wall_b = _timed(lambda: bt.value_and_gradient(thetas), steps, warmup)
        batched_fallbacks = int((bt.last_fallback() != 0).sum())
        bt.close()
        print(f"[bench] batch of {len(probs)}: {name} looped single-problem path", file=sys.stderr,
              flush=True)
        models = [gx.Gpr(k, lik).with_precision(prec).prepare(x, n, d, y) for (k, lik, x, n, d, y) in probs]
        grads = [np.empty(t.size) for t in thetas]

        def loop():
            return [mdl.value_and_gradient_into(t, g_) for mdl, t, g_ in zip(models, thetas, grads)]

        vals, wall_l = _timed(loop, steps, warmup)
        fallbacks = sum(mdl.last_fallback() != 0 for mdl in models)
        for mdl in models:
            mdl.close()
        form = {"batched_ms": wall_b / steps * 1e3, "looped_ms": wall_l / steps * 1e3,
                "speedup": wall_l / wall_b, "failed": sum(e is not None for e in er),
                "batched_fp64_fallbacks": batched_fallbacks,
                "looped_fp64_fallbacks": int(fallbacks),
                "batched_vs_looped_err": {"nlml": err1(nl, vals),
                                          "grad": max(err1(a, b) for a, b in zip(gr, grads))}}
        if ref is None:
            ref = (np.asarray(nl), [g_.copy() for g_ in gr])
        else:
            form["batched_vs_fp64_err"] = {"nlml": err1(nl, ref[0]),
                                           "grad": max(err1(a, b) for a, b in zip(gr, ref[1]))}
            form["looped_vs_fp64_err"] = {"nlml": err1(vals, ref[0]),
                                          "grad": max(err1(a, b) for a, b in zip(grads, ref[1]))}
        out_forms[name] = form
    return out_forms, F


def bench_batch(args, gx):
    forms = tuple(args.forms.split(","))
    out_forms, F = time_batch(gx, forms, args.steps, args.warmup)
    if "mixed" not in out_forms or "fp64" not in out_forms:  # profiling runs of one form
        return {"metric": "batch evals/s, 256 exact ARD-RBF problems (one form)", "forms": out_forms,
                "algorithmic_flop_per_batch": F}
    fb, fm = out_forms["fp64"], out_forms["mixed"]
    ach = F / (fm["batched_ms"] * 1e-3) / 1e12
    ach64 = F / (fb["batched_ms"] * 1e-3) / 1e12
    return {
        "metric": "batch evals/s, 256 exact ARD-RBF problems n in [256, 4096], batched vs looped",
        "value": 1e3 / fb["batched_ms"], "unit": "batch evals/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": fb["batched_ms"], "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (seeded SeededRng, seed 5)",
        "config": {"workload": "batched exact GPR NLML+grad", "problems": 256,
                   "parallelism": "single GPU"},
        "fp64": fb, "mixed": fm, "algorithmic_flop_per_batch": F,
        "roofline": {"bound": "mfma", "kernel": "whole batched FP64 evaluation", "achieved": ach64,
                     "peak": PEAK_FP64, "unit": "TFLOP/s", "frac": ach64 / PEAK_FP64, "traffic": None},
        "roofline_mixed": {"bound": "mfma", "kernel": "whole batched mixed evaluation", "achieved": ach,
                           "peak": PEAK_FP32, "unit": "TFLOP/s", "frac": ach / PEAK_FP32},
    }


def bench_fit(args, gx):
    """Gpr.fit with Lbfgs and R restarts at config 1 (n=1024, d=4, isotropic
    RBF, FP64): every start's L-BFGS evaluated together by the batched path,
    against the same starts fitted one after another."""
    import copy

    from gprx_amd.configs import exact_rbf_small
    from gprx_amd.rng import SeededRng

    wl = exact_rbf_small()
    R, seed = args.restarts, 0
    th0 = wl.theta()
    lo, hi = math.log(1e-5), math.log(1e5)
    r = SeededRng(seed)  # the restart draws gprx_hip_exact_fit makes
    starts = [th0] + [np.array([lo + r.open_unit() * (hi - lo) for _ in th0]) for _ in range(R)]

    def gpr_at(theta):
        spec = copy.deepcopy(wl.kernel())
        spec.set_params(theta[:-1])
        return gx.Gpr(spec, gx.GaussianLikelihood(math.exp(theta[-1])))

    def batched():
        return gpr_at(th0).with_optimizer(gx.Lbfgs().with_restarts(R, seed)).fit(
            wl.x, wl.n, wl.d, wl.y)

    def looped():
        return [gpr_at(t).with_optimizer(gx.Lbfgs()).fit(wl.x, wl.n, wl.d, wl.y) for t in starts]

    print(f"[bench] fit: 1+{R} starts batched", file=sys.stderr, flush=True)
    fb, wall_b = _timed(batched, args.steps, args.warmup)
    best_b = fb.neg_log_marginal_likelihood()
    res_b = fb.fit_result
    fb.close()
    print(f"[bench] fit: 1+{R} starts looped", file=sys.stderr, flush=True)
    fl, wall_l = _timed(looped, args.steps, args.warmup)
    best_l = min(f.neg_log_marginal_likelihood() for f in fl)
    for f in fl:
        f.close()
    # one round of the batched fit: a batched NLML+grad of all starts
    probs = [(gpr_at(t).kernel, gpr_at(t).likelihood, wl.x, wl.n, wl.d, wl.y) for t in starts]
    bt = gx.GprBatch(probs, precision=gx.Precision.FP64)
    _, wall_r = _timed(lambda: bt.value_and_gradient(starts), 5, 1)
    bt.close()
    F = len(starts) * exact_flops(wl.n, wl.d)
    ach = F / (wall_r / 5) / 1e12
    out = {
        "metric": f"L-BFGS starts fitted/s, exact GPR n=1024 d=4 RBF (config 1), 1+{R} starts "
                  "batched vs looped",
        "value": len(starts) / (wall_b / args.steps), "unit": "starts/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": wall_b / args.steps * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (seeded SeededRng, seeds 0 and 1)",
        "config": {"workload": "exact GPR fit, Lbfgs with restarts", "n": wl.n, "d": wl.d,
                   "starts": len(starts), "parallelism": "single GPU"},
        "batched_ms": wall_b / args.steps * 1e3, "looped_ms": wall_l / args.steps * 1e3,
        "speedup": wall_l / wall_b, "best_nlml_batched": best_b, "best_nlml_looped": best_l,
        "best_start": res_b.best_start, "best_start_iterations": res_b.iterations,
        "round_ms": wall_r / 5 * 1e3,
        "roofline": {"bound": "mfma", "kernel": "one round: batched NLML+grad of every start",
                     "achieved": ach, "peak": PEAK_FP64, "unit": "TFLOP/s",
                     "frac": ach / PEAK_FP64, "traffic": None},
    }
    if not args.no_cpu_baseline:
        import oracle

        print("[bench] fit: cpu baseline (oracle objective, first start)", file=sys.stderr,
              flush=True)
        ops = wl.kernel().program()
        span = hi - lo

        def obj(z):
            s_ = 1.0 / (1.0 + np.exp(-z / (span / 4)))
            th = lo + span * s_
            v, gr, _, _ = oracle.exact_value_grad(ops, th, wl.x, wl.n, wl.d, wl.y)
            return v, gr * span * s_ * (1 - s_) / (span / 4)

        u = (th0 - lo) / span
        t0 = time.perf_counter()
        gx.Lbfgs().minimize_unconstrained(obj, np.log(u / (1 - u)) * (span / 4))
        t = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": 1.0 / t, "unit": "starts/s", "cores": oracle.threads(),
                               "kind": "port",
                               "sample": f"the first start's fit on the oracle's objective ({t:.2f} s)"}
    return out


def bench_hessian(args, gx):
    """NLML, gradient and Hessian (FP64) at config 2's kernel and size."""
    from gprx_amd.configs import exact_ard

    wl = exact_ard(n=args.n)
    th = wl.theta()
    p = th.size
    model = gx.Gpr(wl.kernel(), gx.GaussianLikelihood(wl.noise)).prepare(wl.x, wl.n, wl.d, wl.y)
    grad, hess = np.empty(p), np.empty(p * p)
    print(f"[bench] hessian n={wl.n} p={p}", file=sys.stderr, flush=True)
    _, wall = _timed(lambda: model.value_gradient_hessian_into(th, grad, hess), args.steps,
                     args.warmup)
    model.close()
    ms = wall / args.steps * 1e3
    # S_i = (X A_i) X^T per parameter: X A_i with X lower (n^3), then the
    # lower half of (X A_i) X^T with X^T upper (n^3 / 3); plus the evaluation
    F = p * (4.0 / 3.0) * wl.n ** 3 + exact_flops(wl.n, wl.d)
    ach = F / (ms * 1e-3) / 1e12
    H = hess.reshape(p, p)
    return {
        "metric": f"NLML+grad+Hessian evals/s, exact GPR n={wl.n} d={wl.d} Constant x RbfArd (p={p}), FP64",
        "value": 1e3 / ms, "unit": "evals/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic (seeded SeededRng, seed 2)",
        "config": {"workload": "exact GPR Hessian (TrustRegion)", "n": wl.n, "d": wl.d,
                   "params": p, "parallelism": "single GPU"},
        "hessian_symmetric": bool(np.array_equal(H, H.T)),
        "hessian_diag": [float(v) for v in np.diag(H)],
        "algorithmic_flop": F,
        "roofline": {"bound": "mfma", "kernel": "whole Hessian evaluation", "achieved": ach,
                     "peak": PEAK_FP64, "unit": "TFLOP/s", "frac": ach / PEAK_FP64,
                     "traffic": None},
    }


def bench_online(args, gx):
    """OnlineGpr::insert and delete at config 2's kernel: from n = 16384 - K
    points, K insertions inside the 128-row padding (FP64), each O(n^2),
    then K deletions at seeded positions, each O(n^2)."""
    from gprx_amd.configs import exact_ard

    K = args.steps
    wl = exact_ard(n=args.n)
    n0 = wl.n - K - args.warmup
    cols = np.concatenate([wl.x[q * wl.n:q * wl.n + n0] for q in range(wl.d)])
    model = gx.Gpr(wl.kernel(), gx.GaussianLikelihood(wl.noise)).factor(cols, n0, wl.d, wl.y[:n0])

    def pt(i):
        return np.array([wl.x[q * wl.n + i] for q in range(wl.d)])

    print(f"[bench] online: {n0} points, {args.warmup} + {K} insertions", file=sys.stderr,
          flush=True)
    for i in range(n0, n0 + args.warmup):
        model.insert(pt(i), wl.y[i])
    t0 = time.perf_counter()
    for i in range(n0 + args.warmup, wl.n):
        nlml = model.insert(pt(i), wl.y[i])
    wall = time.perf_counter() - t0
    ref = gx.Gpr(wl.kernel(), gx.GaussianLikelihood(wl.noise)).factor(wl.x, wl.n, wl.d, wl.y)
    err = abs(nlml - ref.neg_log_marginal_likelihood()) / abs(ref.neg_log_marginal_likelihood())
    ref.close()
    # OnlineGpr::delete: K deletions at seeded positions, each O(n^2) on X = L^-1
    from gprx_amd.rng import SeededRng

    r = SeededRng(7)
    keep = list(range(wl.n))
    print(f"[bench] online: {K} deletions from {wl.n} points", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    for _ in range(K):
        idx = r.up_to(len(keep) - 1)
        nlml_d = model.delete(idx)
        keep.pop(idx)
    wall_d = time.perf_counter() - t0
    kk = np.array(keep)
    xk = np.concatenate([wl.x[q * wl.n + kk] for q in range(wl.d)])
    ref = gx.Gpr(wl.kernel(), gx.GaussianLikelihood(wl.noise)).factor(xk, kk.size, wl.d, wl.y[kk])
    err_d = abs(nlml_d - ref.neg_log_marginal_likelihood()) / abs(ref.neg_log_marginal_likelihood())
    ref.close()
    model.close()
    ms_d = wall_d / K * 1e3
    ms = wall / K * 1e3
    n_avg = wl.n - K / 2
    # two passes over the lower triangle of X (X k, X^T l)
    bytes_per = 2 * (n_avg * n_avg / 2) * 8
    ach = bytes_per / (ms * 1e-3) / 1e9
    return {
        "metric": f"online insertions/s, exact GPR n~{wl.n} d={wl.d} Constant x RbfArd, FP64",
        "value": 1e3 / ms, "unit": "insertions/s", "n_gpus": 1, "steps": K, "warmup": args.warmup,
        "ms_per_step": ms, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic (seeded SeededRng, seed 2)",
        "config": {"workload": "exact GPR online insertion", "n_final": wl.n, "d": wl.d,
                   "parallelism": "single GPU"},
        "final_nlml_rel_err_vs_fresh_factor": err,
        "roofline": {"bound": "hbm", "kernel": "one insertion (two triangular matrix-vector passes)",
                     "achieved": ach, "peak": PEAK_HBM, "unit": "GB/s", "frac": ach / PEAK_HBM,
                     "traffic": None},
        "delete_ms": ms_d,
        "deletions": K,
        "delete_nlml_rel_err_vs_fresh_factor": err_d,
        # a deletion reads X's lower triangle and writes the shifted one, then
        # reads the new one twice for alpha
        "delete_roofline": {"bound": "hbm", "kernel": "one deletion (X downdate, alpha passes)",
                            "achieved": 4 * (wl.n * wl.n / 2) * 8 / (ms_d * 1e-3) / 1e9,
                            "peak": PEAK_HBM, "unit": "GB/s",
                            "frac": 4 * (wl.n * wl.n / 2) * 8 / (ms_d * 1e-3) / 1e9 / PEAK_HBM},
    }


def bench_persist(args, gx):
    """FittedGpr::save / save_with_factor / LoadedGpr::load at config 2's
    kernel and size (FP64).  The device part of a save with the factor is
    gprx_hip_exact_export_factor: an FP64 factor and inverse, then
    L^T = L^-1 K as one triangular GEMM; a load reads the files and factors
    again on the device."""
    import ctypes as C
    import shutil
    import tempfile

    from gprx_amd import _lib as L
    from gprx_amd.configs import exact_ard

    wl = exact_ard(n=args.n)
    n = wl.n
    model = gx.Gpr(wl.kernel(), gx.GaussianLikelihood(wl.noise)).factor(wl.x, n, wl.d, wl.y)
    l, alpha, jit, err = np.empty(n * n), np.empty(n), C.c_double(), L.GprxError()

    def export():
        rc = L.lib().gprx_hip_exact_export_factor(model._h, L.dptr(l), n, L.dptr(alpha),
                                                   C.byref(jit), C.byref(err))
        if rc != L.OK:
            raise RuntimeError(f"export failed: {rc}")

    print(f"[bench] persist n={n}: factor export x {args.steps}", file=sys.stderr, flush=True)
    _, wall = _timed(export, args.steps, args.warmup)
    ms = wall / args.steps * 1e3
    root = Path(tempfile.mkdtemp(prefix="gprx-persist-"))
    try:
        t0 = time.perf_counter()
        model.save(root / "plain")
        save_ms = (time.perf_counter() - t0) * 1e3
        print("[bench] persist: save_with_factor", file=sys.stderr, flush=True)
        t0 = time.perf_counter()
        model.save_with_factor(root / "factor")
        save_f_ms = (time.perf_counter() - t0) * 1e3
        file_gib = (root / "factor" / "model.safetensors").stat().st_size / 2 ** 30
        print("[bench] persist: load", file=sys.stderr, flush=True)
        t0 = time.perf_counter()
        back = gx.FittedGpr.load(root / "factor")
        load_ms = (time.perf_counter() - t0) * 1e3
        xs = wl.x[:64 * wl.d].copy()
        pa, pb = model.predict(xs, 64, wl.d), back.predict(xs, 64, wl.d)
        pred_err = float(max(np.max(np.abs(pa.mean - pb.mean)), np.max(np.abs(pa.variance - pb.variance))))
        back.close()
    finally:
        shutil.rmtree(root, ignore_errors=True)
    model.close()
    # Cholesky n^3/3, triangular inverse n^3/3, L^-1 K (triangular x full) n^3
    F = (5.0 / 3.0) * float(n) ** 3
    ach = F / (ms * 1e-3) / 1e12
    return {
        "metric": f"factor exports/s (save_with_factor's device part), exact GPR n={n} d={wl.d} "
                  "Constant x RbfArd, FP64",
        "value": 1e3 / ms, "unit": "exports/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic (seeded SeededRng, seed 2)",
        "config": {"workload": "exact GPR save / load", "n": n, "d": wl.d,
                   "parallelism": "single GPU"},
        "save_ms": save_ms, "save_with_factor_ms": save_f_ms, "load_ms": load_ms,
        "factor_file_gib": file_gib, "loaded_prediction_max_abs_diff": pred_err,
        "roofline": {"bound": "mfma", "kernel": "factor export (factor, inverse, L^-1 K; "
                     "including the n^2 readback)", "achieved": ach, "peak": PEAK_FP64,
                     "unit": "TFLOP/s", "frac": ach / PEAK_FP64, "traffic": None},
    }


def bench_copy(args, gx):
    """One 1 GiB device-to-device copy per step (torch's copy kernel): known
    HBM traffic of 1 GiB read + 1 GiB written per launch."""
    import torch

    n = 1 << 27
    src = torch.ones(n, dtype=torch.float64, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(args.warmup):
        dst.copy_(src)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        dst.copy_(src)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    gbs = 2 * n * 8 * args.steps / wall / 1e9
    return {"metric": "device copy GB/s (read + write)", "value": gbs, "unit": "GB/s", "n_gpus": 1,
            "steps": args.steps, "warmup": args.warmup, "ms_per_step": wall / args.steps * 1e3,
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
            "data": "synthetic", "config": {"workload": "copy", "bytes_read_per_launch": n * 8,
                                            "bytes_written_per_launch": n * 8}}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", choices=["exact", "sgpr", "svgp", "batch", "fit", "hessian", "online",
                                           "persist", "copy"],
                    default="exact")
    ap.add_argument("--restarts", type=int, default=16)
    ap.add_argument("--n", type=int, default=16384)
    ap.add_argument("--cpu-warmup", type=int, default=1,
                    help="oracle evaluations discarded before timing (PERF_WARMUP)")
    ap.add_argument("--cpu-reps", type=int, default=1,
                    help="timed oracle evaluations, median reported (PERF_REPS)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--forms", default="fp64,mixed",
                    help="exact and batch workloads: one form alone, e.g. 'fp64' (profiling runs)")
    ap.add_argument("--no-batch", action="store_true",
                    help="skip the 256-problem batch in the exact workload's line")
    ap.add_argument("--no-predict", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: this path runs on one GPU; exact GPR at n=16384 does not shard "
              "across GPUs in this backend; run it on one GPU.", file=sys.stderr)
        sys.exit(2)
    if args.steps is None:
        args.steps = {"exact": 5, "sgpr": 5, "svgp": 2000, "batch": 3, "fit": 1, "hessian": 2,
                      "online": 50, "persist": 3, "copy": 5}[args.workload]

    import gprx_amd as gx

    fn = {"exact": bench_exact, "sgpr": bench_sgpr, "svgp": bench_svgp, "batch": bench_batch,
          "fit": bench_fit, "hessian": bench_hessian, "online": bench_online,
          "persist": bench_persist, "copy": bench_copy}[args.workload]
    print(json.dumps(fn(args, gx)))


if __name__ == "__main__":
    main()
