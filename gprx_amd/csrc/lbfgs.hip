// L-BFGS with restarts (Lbfgs, src/optimizer/lbfgs.rs:19-160; the restart
// driver src/optimizer/mod.rs:182-227; the logit map
// src/optimizer/logit.rs:75-265; the evaluation cache and run comparison
// src/optimizer/adapter.rs:23-330), over exact GPR problems whose starts are
// evaluated together by the batched GPU path, and over a host callback.
//
// The reference calls argmin 0.11's LBFGS with MoreThuenteLineSearch
// (Cargo.lock), which is not in its tree.  The solver here is restated from
// the published methods:
//   - the L-BFGS two-loop recursion with initial scaling gamma = s'y / y'y
//     (Nocedal, Math. Comp. 35, 1980; Nocedal & Wright, Numerical
//     Optimization, 2nd ed., Alg. 7.4), history m, pairs with s'y > 0 kept;
//   - the Moré–Thuente line search (Moré & Thuente, ACM TOMS 20, 1994: the
//     safeguarded step of their dcstep and the interval logic of dcsrch),
//     c1 = 1e-4, c2 = 0.9, relative width tolerance 1e-10, steps in
//     [sqrt(eps), inf), extrapolation bounds 1.1 and 4, first trial step 1,
//     at most 20 evaluations (then the last trial is taken);
//   - stopping when |g| < tolerance, when the cost changes by less than
//     eps between iterations, after max_iterations iterations, or when the
//     direction is not one of descent; the result is the best iterate.
// Every start runs the same solver as its own state machine that asks for
// one point and is told its cost and gradient.  The exact-GPR driver asks
// all live starts, evaluates their points in one batched call and tells
// each its result, so a round costs one batched evaluation however many
// starts there are; the batch shrinks to the live starts as they finish.
#include <cmath>
#include <cstring>
#include <limits>
#include <memory>
#include <vector>

#include "../../include/gprx_hip.h"
#include "logit.hpp"

namespace {

using namespace gprx;

constexpr double kBarrierCost = 1.0e300;  // src/optimizer/adapter.rs:23
constexpr double kEps = 2.220446049250313e-16;
constexpr int kMaxLineEvals = 20;

void put_error(gprx_error* err, int status, const char* msg) {
  if (!err) return;
  std::memset(err, 0, sizeof(*err));
  err->status = status;
  std::strncpy(err->message, msg, sizeof(err->message) - 1);
}

// ------------------------------------------------------- Moré–Thuente
struct LineSearch {
  static constexpr double kFtol = 1e-4, kGtol = 0.9, kXtol = 1e-10;
  static constexpr double kXtrapl = 1.1, kXtrapu = 4.0;
  double stpmin = std::sqrt(kEps), stpmax = std::numeric_limits<double>::infinity();
  bool brackt = false;
  int stage = 1;
  double finit = 0, ginit = 0, gtest = 0, width = 0, width1 = 0;
  double stx = 0, fx = 0, gx = 0, sty = 0, fy = 0, gy = 0, stmin = 0, stmax = 0;
  double stp = 1.0;

  void start(double f0, double g0, double stp0) {
    brackt = false;
    stage = 1;
    finit = f0;
    ginit = g0;
    gtest = kFtol * ginit;
    width = stpmax - stpmin;
    width1 = 2.0 * width;
    stx = 0.0;
    fx = finit;
    gx = ginit;
    sty = 0.0;
    fy = finit;
    gy = ginit;
    stmin = 0.0;
    stp = std::fmin(std::fmax(stp0, stpmin), stpmax);
    stmax = stp + kXtrapu * stp;
  }

  // The safeguarded step of the 1994 paper: a new trial step from the
  // interval end stx (best so far), the other end sty and the trial stp,
  // updating the interval; brackt once a minimizer is bracketed.
  static void dcstep(double& stx, double& fx, double& dx, double& sty, double& fy, double& dy,
                     double& stp, double fp, double dp, bool& brackt, double lo, double hi) {
    const double sgnd = dp * (dx / std::fabs(dx));
    double stpf;
    if (fp > fx) {
      // higher function value: the minimum is bracketed
      const double theta = 3.0 * (fx - fp) / (stp - stx) + dx + dp;
      const double s = std::fmax(std::fabs(theta), std::fmax(std::fabs(dx), std::fabs(dp)));
      double gamma = s * std::sqrt((theta / s) * (theta / s) - (dx / s) * (dp / s));
      if (stp < stx) gamma = -gamma;
      const double p = (gamma - dx) + theta;
      const double q = ((gamma - dx) + gamma) + dp;
      const double r = p / q;
      const double stpc = stx + r * (stp - stx);
      const double stpq = stx + ((dx / ((fx - fp) / (stp - stx) + dx)) / 2.0) * (stp - stx);
      stpf = std::fabs(stpc - stx) < std::fabs(stpq - stx) ? stpc : stpc + (stpq - stpc) / 2.0;
      brackt = true;
    } else if (sgnd < 0.0) {
      // derivatives of opposite sign: bracketed
      const double theta = 3.0 * (fx - fp) / (stp - stx) + dx + dp;
      const double s = std::fmax(std::fabs(theta), std::fmax(std::fabs(dx), std::fabs(dp)));
      double gamma = s * std::sqrt((theta / s) * (theta / s) - (dx / s) * (dp / s));
      if (stp > stx) gamma = -gamma;
      const double p = (gamma - dp) + theta;
      const double q = ((gamma - dp) + gamma) + dx;
      const double r = p / q;
      const double stpc = stp + r * (stx - stp);
      const double stpq = stp + (dp / (dp - dx)) * (stx - stp);
      stpf = std::fabs(stpc - stp) > std::fabs(stpq - stp) ? stpc : stpq;
      brackt = true;
    } else if (std::fabs(dp) < std::fabs(dx)) {
      // same sign, the derivative's magnitude decreases
      const double theta = 3.0 * (fx - fp) / (stp - stx) + dx + dp;
      const double s = std::fmax(std::fabs(theta), std::fmax(std::fabs(dx), std::fabs(dp)));
      double gamma = s * std::sqrt(std::fmax(0.0, (theta / s) * (theta / s) - (dx / s) * (dp / s)));
      if (stp > stx) gamma = -gamma;
      const double p = (gamma - dp) + theta;
      const double q = (gamma + (dx - dp)) + gamma;
      const double r = p / q;
      double stpc;
      if (r < 0.0 && gamma != 0.0)
        stpc = stp + r * (stx - stp);
      else if (stp > stx)
        stpc = hi;
      else
        stpc = lo;
      const double stpq = stp + (dp / (dp - dx)) * (stx - stp);
      if (brackt) {
        stpf = std::fabs(stpc - stp) < std::fabs(stpq - stp) ? stpc : stpq;
        stpf = stp > stx ? std::fmin(stp + 0.66 * (sty - stp), stpf)
                         : std::fmax(stp + 0.66 * (sty - stp), stpf);
      } else {
        stpf = std::fabs(stpc - stp) > std::fabs(stpq - stp) ? stpc : stpq;
        stpf = std::fmax(lo, std::fmin(hi, stpf));
      }
    } else {
      // same sign, the derivative does not decrease
      if (brackt) {
        const double theta = 3.0 * (fp - fy) / (sty - stp) + dy + dp;
        const double s = std::fmax(std::fabs(theta), std::fmax(std::fabs(dy), std::fabs(dp)));
        double gamma = s * std::sqrt((theta / s) * (theta / s) - (dy / s) * (dp / s));
        if (stp > sty) gamma = -gamma;
        const double p = (gamma - dp) + theta;
        const double q = ((gamma - dp) + gamma) + dy;
        const double r = p / q;
        stpf = stp + r * (sty - stp);
      } else {
        stpf = stp > stx ? hi : lo;
      }
    }
    if (fp > fx) {
      sty = stp;
      fy = fp;
      dy = dp;
    } else {
      if (sgnd < 0.0) {
        sty = stx;
        fy = fx;
        dy = dx;
      }
      stx = stp;
      fx = fp;
      dx = dp;
    }
    stp = stpf;
  }

  // Told f and the directional derivative g at stp: true when the search
  // ends at stp (converged or stopped by a safeguard), otherwise stp holds
  // the next trial step.
  bool tell(double f, double g) {
    const double ftest = finit + stp * gtest;
    if (stage == 1 && f <= ftest && g >= 0.0) stage = 2;
    if (brackt && (stp <= stmin || stp >= stmax)) return true;
    if (brackt && stmax - stmin <= kXtol * stmax) return true;
    if (stp == stpmax && f <= ftest && g <= gtest) return true;
    if (stp == stpmin && (f > ftest || g >= gtest)) return true;
    if (f <= ftest && std::fabs(g) <= kGtol * (-ginit)) return true;
    if (stage == 1 && f <= fx && f > ftest) {
      // the modified function psi(stp) = f - stp gtest until it decreases
      double fm = f - stp * gtest, fxm = fx - stx * gtest, fym = fy - sty * gtest;
      double gm = g - gtest, gxm = gx - gtest, gym = gy - gtest;
      dcstep(stx, fxm, gxm, sty, fym, gym, stp, fm, gm, brackt, stmin, stmax);
      fx = fxm + stx * gtest;
      fy = fym + sty * gtest;
      gx = gxm + gtest;
      gy = gym + gtest;
      (void)fm;
    } else {
      dcstep(stx, fx, gx, sty, fy, gy, stp, f, g, brackt, stmin, stmax);
    }
    if (brackt) {
      if (std::fabs(sty - stx) >= 0.66 * width1) stp = stx + 0.5 * (sty - stx);
      width1 = width;
      width = std::fabs(sty - stx);
    }
    if (brackt) {
      stmin = std::fmin(stx, sty);
      stmax = std::fmax(stx, sty);
    } else {
      stmin = stp + kXtrapl * (stp - stx);
      stmax = stp + kXtrapu * (stp - stx);
    }
    stp = std::fmin(std::fmax(stp, stpmin), stpmax);
    if ((brackt && (stp <= stmin || stp >= stmax)) || (brackt && stmax - stmin <= kXtol * stmax))
      stp = stx;
    return false;
  }
};

double dot(const std::vector<double>& a, const std::vector<double>& b) {
  double s = 0.0;
  for (size_t i = 0; i < a.size(); ++i) s += a[i] * b[i];
  return s;
}

// ------------------------------------------------------------- L-BFGS run
struct LbfgsRun {
  enum Phase { kInit, kSearch, kDone };
  int n, m;
  double tol_grad;
  uint64_t max_iter;
  Phase phase = kInit;
  uint64_t iter = 0;
  std::vector<double> x, g, d, xt;
  double f = 0.0;
  std::vector<std::vector<double>> S, Y;  // oldest first
  std::vector<double> best_x;
  double best_f = std::numeric_limits<double>::infinity();
  LineSearch ls;
  int line_evals = 0;

  LbfgsRun(const double* x0, int n_, const gprx_lbfgs_options& o)
      : n(n_), m(o.history_size), tol_grad(o.tolerance), max_iter(o.max_iterations),
        x(x0, x0 + n_), g(n_), d(n_), xt(x0, x0 + n_), best_x(x0, x0 + n_) {}

  bool done() const { return phase == kDone; }
  const double* ask() const { return xt.data(); }

  void keep_best(const std::vector<double>& p, double fv) {
    if (fv < best_f) {
      best_f = fv;
      best_x = p;
    }
  }

  // The two-loop recursion: d = -H g.
  void direction() {
    const int k = (int)S.size();
    std::vector<double> q = g, alpha(k), rho(k);
    for (int i = k - 1; i >= 0; --i) {
      rho[i] = 1.0 / dot(Y[i], S[i]);
      alpha[i] = rho[i] * dot(S[i], q);
      for (int j = 0; j < n; ++j) q[j] -= alpha[i] * Y[i][j];
    }
    const double gamma = k > 0 ? dot(S[k - 1], Y[k - 1]) / dot(Y[k - 1], Y[k - 1]) : 1.0;
    for (int j = 0; j < n; ++j) q[j] *= gamma;
    for (int i = 0; i < k; ++i) {
      const double beta = rho[i] * dot(Y[i], q);
      for (int j = 0; j < n; ++j) q[j] += S[i][j] * (alpha[i] - beta);
    }
    for (int j = 0; j < n; ++j) d[j] = -q[j];
  }

  bool converged_at_start_of_iteration(double prev_f) const {
    if (std::sqrt(dot(g, g)) < tol_grad) return true;
    if (std::fabs(prev_f - f) < kEps) return true;
    return iter >= max_iter;
  }

  void begin_iteration() {
    direction();
    const double gd = dot(g, d);
    if (!(gd < 0.0)) {
      // no descent along d (only with kept pairs from inexact gradients):
      // the run ends at its best iterate
      phase = kDone;
      return;
    }
    ls.start(f, gd, 1.0);
    line_evals = 0;
    for (int j = 0; j < n; ++j) xt[j] = x[j] + ls.stp * d[j];
    phase = kSearch;
  }

  void tell(double fv, const double* gv) {
    if (phase == kInit) {
      f = fv;
      g.assign(gv, gv + n);
      keep_best(x, f);
      if (converged_at_start_of_iteration(std::numeric_limits<double>::infinity())) {
        phase = kDone;
        return;
      }
      begin_iteration();
      return;
    }
    if (phase != kSearch) return;
    ++line_evals;
    double gd = 0.0;
    for (int j = 0; j < n; ++j) gd += gv[j] * d[j];
    const bool end = ls.tell(fv, gd) || line_evals >= kMaxLineEvals;
    if (!end) {
      for (int j = 0; j < n; ++j) xt[j] = x[j] + ls.stp * d[j];
      return;
    }
    // accept the last trial point: s = x_new - x, y = g_new - g
    std::vector<double> s(n), y(n);
    for (int j = 0; j < n; ++j) {
      s[j] = xt[j] - x[j];
      y[j] = gv[j] - g[j];
    }
    // A pair without positive curvature (s'y <= 0: a line search stopped
    // before the curvature condition, or gradients good only to f32) would
    // leave the implicit inverse Hessian indefinite: it is not kept.
    if (dot(s, y) > 0.0) {
      if ((int)S.size() >= m) {
        S.erase(S.begin());
        Y.erase(Y.begin());
      }
      S.push_back(std::move(s));
      Y.push_back(std::move(y));
    }
    const double prev_f = f;
    x = xt;
    f = fv;
    g.assign(gv, gv + n);
    ++iter;
    keep_best(x, f);
    if (converged_at_start_of_iteration(prev_f)) {
      phase = kDone;
      return;
    }
    begin_iteration();
  }
};

// One evaluation's result as the reference's evaluation cache reports it: a
// failed or non-finite evaluation costs BARRIER_COST with a zero gradient
// (src/optimizer/adapter.rs:116-150).
void barrier_if_bad(int status, double* f, double* g, int n) {
  bool ok = status == GPRX_OK && std::isfinite(*f);
  for (int j = 0; ok && j < n; ++j) ok = std::isfinite(g[j]);
  if (!ok) {
    *f = kBarrierCost;
    for (int j = 0; j < n; ++j) g[j] = 0.0;
  }
}

bool valid_options(const gprx_lbfgs_options* o, gprx_error* err) {
  if (!o) return true;
  if (!std::isfinite(o->tolerance) || o->tolerance < 0.0) {
    put_error(err, GPRX_INVALID_CONFIG, "L-BFGS gradient tolerance must be finite and >= 0");
    return false;
  }
  if (o->history_size < 1) {
    put_error(err, GPRX_INVALID_CONFIG, "L-BFGS history size must be at least 1");
    return false;
  }
  return true;
}

gprx_lbfgs_options options_or_default(const gprx_lbfgs_options* o) {
  gprx_lbfgs_options d;
  gprx_lbfgs_default_options(&d);
  return o ? *o : d;
}

}  // namespace

extern "C" {

void gprx_lbfgs_default_options(gprx_lbfgs_options* opt) {
  if (!opt) return;
  opt->max_iterations = 100;
  opt->tolerance = std::sqrt(kEps);
  opt->history_size = 10;
  opt->restarts = 0;
  opt->seed = 0;
}

int gprx_lbfgs_minimize(gprx_objective_fn fn, void* ctx, const double* x0, size_t n,
                        const gprx_lbfgs_options* opt, double* x_out, double* value,
                        uint64_t* iterations, gprx_error* err) {
  if (!fn || !x0 || !x_out || n == 0) {
    put_error(err, GPRX_LENGTH_MISMATCH, "objective, start and output are required");
    return GPRX_LENGTH_MISMATCH;
  }
  if (!valid_options(opt, err)) return GPRX_INVALID_CONFIG;
  const gprx_lbfgs_options o = options_or_default(opt);
  LbfgsRun run(x0, (int)n, o);
  std::vector<double> g(n);
  while (!run.done()) {
    double f = 0.0;
    const int st = fn(ctx, run.ask(), n, &f, g.data());
    barrier_if_bad(st, &f, g.data(), (int)n);
    run.tell(f, g.data());
  }
  std::memcpy(x_out, run.best_x.data(), n * sizeof(double));
  if (value) *value = run.best_f;
  if (iterations) *iterations = run.iter;
  if (!(run.best_f < kBarrierCost)) {
    put_error(err, GPRX_NOT_CONVERGED, "L-BFGS found no point with a finite cost");
    return GPRX_NOT_CONVERGED;
  }
  if (err) put_error(err, GPRX_OK, "");
  return GPRX_OK;
}

int gprx_hip_exact_fit(const double* x, size_t n, size_t d, const double* y,
                       const gprx_kernel_program* kernel, const double* log_theta, size_t p,
                       const double* bounds, int32_t precision, const gprx_jitter_policy* jitter,
                       const gprx_lbfgs_options* opt, double* best_log_theta, double* best_value,
                       uint64_t* iterations, int32_t* best_start, gprx_error* err) {
  if (!log_theta || !best_log_theta || p == 0) {
    put_error(err, GPRX_LENGTH_MISMATCH, "log_theta and best_log_theta are required");
    return GPRX_LENGTH_MISMATCH;
  }
  if (!valid_options(opt, err)) return GPRX_INVALID_CONFIG;
  const gprx_lbfgs_options o = options_or_default(opt);
  const int np = (int)p;
  std::vector<Interval> iv(np);
  for (int i = 0; i < np; ++i) {
    iv[i] = bounds ? Interval{bounds[2 * i], bounds[2 * i + 1]} : Interval{1e-5, 1e5};
    if (!(iv[i].lo < iv[i].hi) || !std::isfinite(iv[i].lo) || !std::isfinite(iv[i].hi)) {
      put_error(err, GPRX_INVALID_CONFIG, "every bound must be a finite interval lo < hi");
      return GPRX_INVALID_CONFIG;
    }
  }
  // Starts in z: the given point, then `restarts` draws (mod.rs:200-226).
  const int k = 1 + (int)o.restarts;
  std::vector<std::vector<double>> z0(k, std::vector<double>(np));
  for (int i = 0; i < np; ++i) {
    const double user = iv[i].positive() ? std::exp(log_theta[i]) : log_theta[i];
    if (!user_to_z(user, iv[i], &z0[0][i])) {
      put_error(err, GPRX_INVALID_INTERVAL, "a starting parameter lies outside its bounds");
      return GPRX_INVALID_INTERVAL;
    }
  }
  Rng rng(o.seed);
  for (int r = 1; r < k; ++r)
    for (int i = 0; i < np; ++i) user_to_z(sample_user(rng, iv[i]), iv[i], &z0[r][i]);

  std::vector<LbfgsRun> runs;
  runs.reserve(k);
  for (int r = 0; r < k; ++r) runs.emplace_back(z0[r].data(), np, o);
  // The batch holds one problem (the same data) per start in `members`.
  // When the live starts fall to half the batch, it is made again over
  // them alone, with the launch schedule the first batch tuned, so the
  // long tail of a few slow starts does not pay for the finished ones.
  std::vector<int> members;
  gprx_hip_model* model = nullptr;
  std::unique_ptr<gprx_hip_model, void (*)(gprx_hip_model*)> hold(nullptr, gprx_hip_destroy);
  int sched_policy = -1, sched_small = 0;
  auto make_batch = [&](const std::vector<int>& who) -> int {
    std::vector<gprx_problem> probs(who.size(), gprx_problem{x, n, d, y, kernel, log_theta, p});
    gprx_hip_model* mnew = nullptr;
    int rc2 = gprx_hip_batch_create(probs.data(), probs.size(), precision, &mnew, err);
    if (rc2 != GPRX_OK) return rc2;
    hold.reset(mnew);
    model = mnew;
    members = who;
    if (sched_policy >= 0) {
      rc2 = gprx_hip_batch_set_schedule(model, sched_policy, sched_small, err);
      if (rc2 != GPRX_OK) return rc2;
    }
    return GPRX_OK;
  };
  std::vector<int> all(k);
  for (int r = 0; r < k; ++r) all[r] = r;
  int rc = make_batch(all);
  if (rc != GPRX_OK) return rc;
  std::vector<std::vector<double>> th(k, std::vector<double>(np)), gr(k, std::vector<double>(np));
  std::vector<const double*> thp(k);
  std::vector<double*> grp(k);
  std::vector<double> nlml(k);
  std::vector<gprx_error> errs(k);
  // The batch factors every point without jitter.  A point whose factor
  // fails is evaluated again alone under the model's jitter policy, as the
  // reference's fit evaluates every point (src/gpr/exact_fit.rs:280-283).
  const bool retry = jitter && jitter->max_retries > 0;
  std::unique_ptr<gprx_hip_model, void (*)(gprx_hip_model*)> alone(nullptr, gprx_hip_destroy);
  auto eval_alone = [&](int s) -> int {
    if (!alone) {
      gprx_hip_model* m1 = nullptr;
      const int rc1 = gprx_hip_exact_create(x, n, d, y, kernel, thp[s], p, precision, jitter, &m1,
                                            &errs[s]);
      if (rc1 != GPRX_OK) return rc1;
      alone.reset(m1);
    }
    return gprx_hip_exact_value_and_grad(alone.get(), thp[s], p, &nlml[s], grp[s], &errs[s]);
  };
  // The status of each start's first evaluation: the fit's error when the
  // first start finds no finite point (src/optimizer/mod.rs:210-214 with
  // best_value's re-evaluation, src/optimizer/adapter.rs:266-278).
  std::vector<gprx_error> start_err(k);
  std::vector<char> started(k, 0);
  for (;;) {
    std::vector<int> live;
    for (int r = 0; r < k; ++r)
      if (!runs[r].done()) live.push_back(r);
    if (live.empty()) break;
    if (members.size() > 1 && live.size() * 2 <= members.size()) {
      if (sched_policy < 0) {
        double tuned = 0.0;
        (void)gprx_hip_batch_schedule(model, &sched_policy, &sched_small, &tuned);
      }
      rc = make_batch(live);
      if (rc != GPRX_OK) return rc;
    }
    const int kb = (int)members.size();
    for (int s = 0; s < kb; ++s) {
      // a member that has finished is evaluated again at its best point
      // (the result unused): a batch evaluates all of its problems
      const LbfgsRun& run = runs[members[s]];
      const double* z = run.done() ? run.best_x.data() : run.ask();
      for (int i = 0; i < np; ++i) th[s][i] = z_to_param(z[i], iv[i]);
      thp[s] = th[s].data();
      grp[s] = gr[s].data();
    }
    rc = gprx_hip_batch_value_and_grad(model, thp.data(), nlml.data(), grp.data(), errs.data());
    if (rc != GPRX_OK) {
      put_error(err, rc, errs[0].message[0] ? errs[0].message : "batched evaluation failed");
      return rc;
    }
    for (int s = 0; s < kb; ++s) {
      LbfgsRun& run = runs[members[s]];
      if (run.done()) continue;
      if (retry && errs[s].status == GPRX_CHOLESKY_FAILED) {
        errs[s].status = eval_alone(s);
        if (errs[s].status == GPRX_DEVICE_ERROR) {
          if (err) *err = errs[s];
          return GPRX_DEVICE_ERROR;
        }
      }
      if (!started[members[s]]) {
        started[members[s]] = 1;
        start_err[members[s]] = errs[s];
      }
      const double* z = run.ask();
      for (int i = 0; i < np; ++i) gr[s][i] *= dparam_dz(z[i], iv[i]);
      double f = nlml[s];
      barrier_if_bad(errs[s].status, &f, gr[s].data(), np);
      run.tell(f, gr[s].data());
    }
  }
  // keep_better over the runs in order; the first start's failure is the
  // fit's (mod.rs:210-214), a failed restart is skipped
  int best = -1;
  for (int r = 0; r < k; ++r) {
    const bool ok = runs[r].best_f < kBarrierCost && std::isfinite(runs[r].best_f);
    if (!ok) {
      if (r == 0) {
        if (start_err[0].status != GPRX_OK) {
          if (err) *err = start_err[0];
          return start_err[0].status;
        }
        put_error(err, GPRX_NOT_CONVERGED, "the first start found no point with a finite cost");
        return GPRX_NOT_CONVERGED;
      }
      continue;
    }
    if (best < 0 || runs[r].best_f < runs[best].best_f) best = r;
  }
  for (int i = 0; i < np; ++i) best_log_theta[i] = z_to_param(runs[best].best_x[i], iv[i]);
  if (best_value) *best_value = runs[best].best_f;
  if (iterations) *iterations = runs[best].iter;
  if (best_start) *best_start = best;
  if (err) put_error(err, GPRX_OK, "");
  return GPRX_OK;
}

}  // extern "C"
