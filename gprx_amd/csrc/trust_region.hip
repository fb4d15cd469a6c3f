// Trust-region minimization with the analytic Hessian (TrustRegion,
// src/optimizer/trust_region.rs: radius 1, largest radius 100, 100
// iterations, stop when |g| < sqrt(eps), Steihaug subproblem of at most
// 4 p + 10 CG steps; with restarts as src/optimizer/mod.rs:182-227), over a
// host callback and over an exact GPR model whose Hessian the GPU forms
// (gprx_hip_exact_hessian), through the logit map of the parameter bounds
// with its second-order term (src/optimizer/logit.rs:43-73,160-190).
//
// The reference calls argmin 0.11's TrustRegion with its Steihaug
// subproblem (Cargo.lock), which is not in its tree.  Restated here from the
// published methods (Nocedal & Wright, Numerical Optimization, 2nd ed.):
//   - Alg. 4.1: rho = actual / predicted decrease; radius / 4 when
//     rho < 1/4, doubled (up to the largest) when rho > 3/4 and the step
//     reached the boundary; the step is taken when rho > 1/8;
//   - Alg. 7.2 (CG-Steihaug): CG on B p = -g from p = 0, stopping at the
//     boundary along a direction of non-positive curvature or a step that
//     leaves the region, or when the residual falls below 1e-10.
// Every point is evaluated with its gradient and Hessian, as the
// reference's evaluation cache does; a failed point costs 1e300.
#include <cmath>
#include <cstring>
#include <limits>
#include <vector>

#include "../../include/gprx_hip.h"
#include "logit.hpp"

namespace {

using namespace gprx;

constexpr double kBarrier = 1.0e300;
constexpr double kEps = 2.220446049250313e-16;

void put_error(gprx_error* err, int status, const char* msg) {
  if (!err) return;
  std::memset(err, 0, sizeof(*err));
  err->status = status;
  std::strncpy(err->message, msg, sizeof(err->message) - 1);
}

double dot(const double* a, const double* b, int n) {
  double s = 0.0;
  for (int i = 0; i < n; ++i) s += a[i] * b[i];
  return s;
}

// Value, gradient and Hessian (row-major) at x; false when the point failed.
struct Objective2 {
  virtual ~Objective2() = default;
  virtual bool eval(const double* x, double* f, double* g, double* H) = 0;
};

// tau >= 0 with |p + tau d| = radius
double to_boundary(const double* p, const double* d, int n, double radius) {
  const double a = dot(d, d, n), b = 2.0 * dot(p, d, n), c = dot(p, p, n) - radius * radius;
  const double disc = std::sqrt(std::fmax(0.0, b * b - 4.0 * a * c));
  return (-b + disc) / (2.0 * a);
}

// CG-Steihaug: approximate minimizer of g'p + p'Bp/2 over |p| <= radius.
void steihaug(const double* g, const double* B, int n, double radius, double* p) {
  std::vector<double> r(g, g + n), d(n), Bd(n), pn(n);
  for (int i = 0; i < n; ++i) {
    p[i] = 0.0;
    d[i] = -r[i];
  }
  if (std::sqrt(dot(r.data(), r.data(), n)) < 1e-10) return;
  const int max_it = 4 * n + 10;
  for (int it = 0; it < max_it; ++it) {
    for (int i = 0; i < n; ++i) Bd[i] = dot(B + (size_t)i * n, d.data(), n);
    const double dBd = dot(d.data(), Bd.data(), n);
    if (dBd <= 0.0) {
      const double tau = to_boundary(p, d.data(), n, radius);
      for (int i = 0; i < n; ++i) p[i] += tau * d[i];
      return;
    }
    const double rr = dot(r.data(), r.data(), n);
    const double alpha = rr / dBd;
    for (int i = 0; i < n; ++i) pn[i] = p[i] + alpha * d[i];
    if (std::sqrt(dot(pn.data(), pn.data(), n)) >= radius) {
      const double tau = to_boundary(p, d.data(), n, radius);
      for (int i = 0; i < n; ++i) p[i] += tau * d[i];
      return;
    }
    for (int i = 0; i < n; ++i) {
      p[i] = pn[i];
      r[i] += alpha * Bd[i];
    }
    const double rn = dot(r.data(), r.data(), n);
    if (std::sqrt(rn) < 1e-10) return;
    const double beta = rn / rr;
    for (int i = 0; i < n; ++i) d[i] = -r[i] + beta * d[i];
  }
}

struct TrResult {
  std::vector<double> x;
  double f = std::numeric_limits<double>::infinity();
  uint64_t iterations = 0;
  bool start_failed = false;
};

TrResult trust_region(Objective2& obj, const double* x0, int n, const gprx_tr_options& o) {
  TrResult res;
  std::vector<double> x(x0, x0 + n), g(n), H((size_t)n * n), xt(n), gt(n), Ht((size_t)n * n), p(n);
  double f = 0.0;
  if (!obj.eval(x.data(), &f, g.data(), H.data())) {
    res.start_failed = true;
    res.x = x;
    return res;
  }
  res.x = x;
  res.f = f;
  double radius = o.initial_radius;
  uint64_t it = 0;
  for (; it < o.max_iterations; ++it) {
    if (std::sqrt(dot(g.data(), g.data(), n)) < o.tolerance) break;
    steihaug(g.data(), H.data(), n, radius, p.data());
    std::vector<double> Bp(n);
    for (int i = 0; i < n; ++i) Bp[i] = dot(H.data() + (size_t)i * n, p.data(), n);
    const double pred = -(dot(g.data(), p.data(), n) + 0.5 * dot(p.data(), Bp.data(), n));
    for (int i = 0; i < n; ++i) xt[i] = x[i] + p[i];
    double ft = kBarrier;
    const bool ok = obj.eval(xt.data(), &ft, gt.data(), Ht.data());
    if (!ok) ft = kBarrier;
    const double rho = pred > 0.0 ? (f - ft) / pred : -1.0;
    const double pnorm = std::sqrt(dot(p.data(), p.data(), n));
    if (rho < 0.25)
      radius = 0.25 * radius;
    else if (rho > 0.75 && std::fabs(pnorm - radius) <= 1e-12 * radius)
      radius = std::fmin(2.0 * radius, o.max_radius);
    if (ok && rho > 0.125) {
      x = xt;
      f = ft;
      g = gt;
      H = Ht;
      if (f < res.f) {
        res.f = f;
        res.x = x;
      }
    }
    if (!(radius > 1e-300)) {
      ++it;
      break;
    }
  }
  res.iterations = it;
  return res;
}

struct CallbackObjective : Objective2 {
  gprx_objective2_fn fn;
  void* ctx;
  int n;
  bool eval(const double* x, double* f, double* g, double* H) override {
    const int st = fn(ctx, x, (size_t)n, f, g, H);
    bool ok = st == 0 && std::isfinite(*f);
    for (int i = 0; ok && i < n; ++i) ok = std::isfinite(g[i]);
    for (int i = 0; ok && i < n * n; ++i) ok = std::isfinite(H[i]);
    return ok;
  }
};

// The exact model's NLML, gradient and Hessian in logit coordinates:
// H_z = J H J + diag(g * d2theta/dz2), g_z = J g.
struct ExactLogitObjective : Objective2 {
  gprx_hip_model* model;
  std::vector<Interval> iv;
  int n;
  // evaluations since the run started, and the status of its first (the
  // start), which src/optimizer/trust_region.rs:166-168 returns as the run's
  // error when the start cannot be evaluated
  int calls = 0;
  int start_rc = GPRX_OK;
  gprx_error start_err{};
  bool eval(const double* z, double* f, double* g, double* H) override {
    std::vector<double> th(n), gt(n), Ht((size_t)n * n);
    for (int i = 0; i < n; ++i) th[i] = z_to_param(z[i], iv[i]);
    gprx_error e{};
    const int rc = gprx_hip_exact_hessian(model, th.data(), (size_t)n, f, gt.data(), Ht.data(), &e);
    if (++calls == 1) {
      start_rc = rc;
      start_err = e;
    }
    if (rc != GPRX_OK) return false;
    if (!std::isfinite(*f)) return false;
    for (int i = 0; i < n; ++i) {
      const double ji = dparam_dz(z[i], iv[i]);
      g[i] = gt[i] * ji;
      for (int j = 0; j < n; ++j) H[(size_t)i * n + j] = Ht[(size_t)i * n + j] * ji * dparam_dz(z[j], iv[j]);
      H[(size_t)i * n + i] += gt[i] * d2param_dz2(z[i], iv[i]);
    }
    for (int i = 0; i < n; ++i)
      if (!std::isfinite(g[i])) return false;
    return true;
  }
};

bool valid(const gprx_tr_options& o, gprx_error* err) {
  if (!std::isfinite(o.tolerance) || o.tolerance < 0.0) {
    put_error(err, GPRX_INVALID_CONFIG, "trust-region gradient tolerance must be finite and >= 0");
    return false;
  }
  if (!(std::isfinite(o.initial_radius) && std::isfinite(o.max_radius) && o.initial_radius > 0.0 &&
        o.initial_radius <= o.max_radius)) {
    put_error(err, GPRX_INVALID_CONFIG, "trust-region radii need 0 < initial <= largest");
    return false;
  }
  return true;
}

}  // namespace

extern "C" {

void gprx_tr_default_options(gprx_tr_options* opt) {
  if (!opt) return;
  opt->max_iterations = 100;
  opt->tolerance = std::sqrt(kEps);
  opt->initial_radius = 1.0;
  opt->max_radius = 100.0;
  opt->restarts = 0;
  opt->seed = 0;
}

int gprx_trust_region_minimize(gprx_objective2_fn fn, void* ctx, const double* x0, size_t n,
                               const gprx_tr_options* opt, double* x_out, double* value,
                               uint64_t* iterations, gprx_error* err) {
  if (!fn || !x0 || !x_out || n == 0) {
    put_error(err, GPRX_LENGTH_MISMATCH, "objective, start and output are required");
    return GPRX_LENGTH_MISMATCH;
  }
  gprx_tr_options o;
  gprx_tr_default_options(&o);
  if (opt) o = *opt;
  if (!valid(o, err)) return GPRX_INVALID_CONFIG;
  CallbackObjective obj;
  obj.fn = fn;
  obj.ctx = ctx;
  obj.n = (int)n;
  TrResult r = trust_region(obj, x0, (int)n, o);
  std::memcpy(x_out, r.x.data(), n * sizeof(double));
  if (value) *value = r.f;
  if (iterations) *iterations = r.iterations;
  if (r.start_failed || !(r.f < kBarrier)) {
    put_error(err, GPRX_NOT_CONVERGED, "the start could not be evaluated");
    return GPRX_NOT_CONVERGED;
  }
  if (err) put_error(err, GPRX_OK, "");
  return GPRX_OK;
}

int gprx_hip_exact_fit_trust_region(gprx_hip_model* model, const double* log_theta, size_t p,
                                    const double* bounds, const gprx_tr_options* opt,
                                    double* best_log_theta, double* best_value,
                                    uint64_t* iterations, int32_t* best_start, gprx_error* err) {
  if (!model || !log_theta || !best_log_theta || p == 0) {
    put_error(err, GPRX_LENGTH_MISMATCH, "model, log_theta and best_log_theta are required");
    return GPRX_LENGTH_MISMATCH;
  }
  gprx_tr_options o;
  gprx_tr_default_options(&o);
  if (opt) o = *opt;
  if (!valid(o, err)) return GPRX_INVALID_CONFIG;
  const int np = (int)p;
  ExactLogitObjective obj;
  obj.model = model;
  obj.n = np;
  obj.iv.resize(np);
  for (int i = 0; i < np; ++i) {
    obj.iv[i] = bounds ? Interval{bounds[2 * i], bounds[2 * i + 1]} : Interval{1e-5, 1e5};
    if (!(obj.iv[i].lo < obj.iv[i].hi) || !std::isfinite(obj.iv[i].lo) || !std::isfinite(obj.iv[i].hi)) {
      put_error(err, GPRX_INVALID_CONFIG, "every bound must be a finite interval lo < hi");
      return GPRX_INVALID_CONFIG;
    }
  }
  std::vector<double> z0(np);
  for (int i = 0; i < np; ++i) {
    const double user = obj.iv[i].positive() ? std::exp(log_theta[i]) : log_theta[i];
    if (!user_to_z(user, obj.iv[i], &z0[i])) {
      put_error(err, GPRX_INVALID_INTERVAL, "a starting parameter lies outside its bounds");
      return GPRX_INVALID_INTERVAL;
    }
  }
  // the given start, then restart draws (src/optimizer/mod.rs:200-226); the
  // first start's failure fails the fit, a failed restart is skipped
  TrResult best;
  int best_i = -1;
  Rng rng(o.seed);
  for (uint32_t r = 0; r <= o.restarts; ++r) {
    if (r > 0)
      for (int i = 0; i < np; ++i) user_to_z(sample_user(rng, obj.iv[i]), obj.iv[i], &z0[i]);
    obj.calls = 0;
    obj.start_rc = GPRX_OK;
    TrResult t = trust_region(obj, z0.data(), np, o);
    const bool ok = !t.start_failed && t.f < kBarrier && std::isfinite(t.f);
    if (!ok) {
      if (r == 0) {
        // the start's own error (CholeskyFailed, InvalidHyperparameter, ...)
        if (t.start_failed && obj.start_rc != GPRX_OK) {
          if (err) *err = obj.start_err;
          return obj.start_rc;
        }
        put_error(err, GPRX_NOT_CONVERGED, "the first start could not be evaluated");
        return GPRX_NOT_CONVERGED;
      }
      continue;
    }
    if (best_i < 0 || t.f < best.f) {
      best = t;
      best_i = (int)r;
    }
  }
  for (int i = 0; i < np; ++i) best_log_theta[i] = z_to_param(best.x[i], obj.iv[i]);
  // leave the model at the best point
  double v = 0.0;
  gprx_error e{};
  const int rc = gprx_hip_exact_value_and_grad(model, best_log_theta, p, &v, nullptr, &e);
  if (rc != GPRX_OK) {
    if (err) *err = e;
    return rc;
  }
  if (best_value) *best_value = best.f;
  if (iterations) *iterations = best.iterations;
  if (best_start) *best_start = best_i;
  if (err) put_error(err, GPRX_OK, "");
  return GPRX_OK;
}

}  // extern "C"
