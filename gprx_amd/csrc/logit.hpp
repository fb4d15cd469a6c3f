// The logit map of the optimizers (src/optimizer/logit.rs:75-265): each
// parameter's interval in user units, log-uniform on positive intervals
// (where the model parameter is log theta) and uniform otherwise, with the
// coordinate scaled so d param / dz = 1 at the interval's middle.
#pragma once

#include <cmath>

#include "rng.hpp"

namespace gprx {

struct Interval {
  double lo, hi;
  bool positive() const { return lo > 0.0; }
  double span() const { return positive() ? std::log(hi) - std::log(lo) : hi - lo; }
  double scale() const { return span() / 4.0; }
  bool contains(double v) const { return std::isfinite(v) && v > lo && v < hi; }
};

inline double sigmoid(double z) {
  if (z >= 0.0) return 1.0 / (1.0 + std::exp(-z));
  const double e = std::exp(z);
  return e / (1.0 + e);
}

// The model parameter of z: log theta on a positive interval, theta otherwise.
inline double z_to_param(double z, const Interval& iv) {
  const double u = sigmoid(z / iv.scale());
  return iv.positive() ? std::log(iv.lo) + u * (std::log(iv.hi) - std::log(iv.lo))
                       : iv.lo + u * (iv.hi - iv.lo);
}

// d param / dz and d2 param / dz2
inline double dparam_dz(double z, const Interval& iv) {
  const double s = sigmoid(z / iv.scale());
  return iv.span() * s * (1.0 - s) / iv.scale();
}
inline double d2param_dz2(double z, const Interval& iv) {
  const double s = sigmoid(z / iv.scale());
  return iv.span() * s * (1.0 - s) * (1.0 - 2.0 * s) / (iv.scale() * iv.scale());
}

// z of a user value x strictly inside the interval (false outside).
inline bool user_to_z(double x, const Interval& iv, double* z) {
  if (!iv.contains(x)) return false;
  constexpr double kEps = 2.220446049250313e-16;
  double u = iv.positive() ? (std::log(x) - std::log(iv.lo)) / (std::log(iv.hi) - std::log(iv.lo))
                           : (x - iv.lo) / (iv.hi - iv.lo);
  u = std::fmin(std::fmax(u, kEps), 1.0 - kEps);
  *z = std::log(u / (1.0 - u)) * iv.scale();
  return true;
}

// A restart draw: log-uniform (positive intervals) or uniform, strictly inside.
inline double sample_user(Rng& rng, const Interval& iv) {
  const double u = rng.open_unit();
  return iv.positive() ? std::exp(std::log(iv.lo) + u * (std::log(iv.hi) - std::log(iv.lo)))
                       : iv.lo + u * (iv.hi - iv.lo);
}

}  // namespace gprx
