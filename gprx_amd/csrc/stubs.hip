// Entry points whose kernels are not built yet: they return GPRX_UNSUPPORTED.
#include "model.hpp"

extern "C" {
int gprx_hip_sgpr_create(const double*, size_t, size_t, const double*, const double*, size_t,
                         const gprx_kernel_program*, const double*, size_t, int32_t,
                         gprx_hip_model** out, gprx_error* err) {
  if (out) *out = nullptr;
  gprx::set_error(err, GPRX_UNSUPPORTED, "sgpr not built");
  return GPRX_UNSUPPORTED;
}
int gprx_hip_sgpr_value_and_grad(gprx_hip_model*, const double*, size_t, double*, double*,
                                 gprx_error* err) {
  gprx::set_error(err, GPRX_UNSUPPORTED, "sgpr not built");
  return GPRX_UNSUPPORTED;
}
int gprx_hip_sgpr_predict(gprx_hip_model*, const double*, size_t, size_t, int32_t, double*,
                          double*, gprx_error* err) {
  gprx::set_error(err, GPRX_UNSUPPORTED, "sgpr not built");
  return GPRX_UNSUPPORTED;
}
int gprx_hip_svgp_create(const double*, size_t, size_t, const double*, const double*, size_t,
                         const gprx_kernel_program*, const double*, size_t, int32_t,
                         gprx_hip_model** out, gprx_error* err) {
  if (out) *out = nullptr;
  gprx::set_error(err, GPRX_UNSUPPORTED, "svgp not built");
  return GPRX_UNSUPPORTED;
}
size_t gprx_hip_svgp_num_params(const gprx_hip_model*) { return 0; }
int gprx_hip_svgp_value_and_grad(gprx_hip_model*, const double*, size_t, const uint64_t*, size_t,
                                 double*, double*, gprx_error* err) {
  gprx::set_error(err, GPRX_UNSUPPORTED, "svgp not built");
  return GPRX_UNSUPPORTED;
}
int gprx_hip_svgp_adam(gprx_hip_model*, double*, size_t, size_t, size_t, double, double, double,
                       double, uint64_t, double*, gprx_error* err) {
  gprx::set_error(err, GPRX_UNSUPPORTED, "svgp not built");
  return GPRX_UNSUPPORTED;
}
int gprx_hip_svgp_predict(gprx_hip_model*, const double*, size_t, const double*, size_t, size_t,
                          int32_t, double*, double*, gprx_error* err) {
  gprx::set_error(err, GPRX_UNSUPPORTED, "svgp not built");
  return GPRX_UNSUPPORTED;
}
int gprx_hip_batch_create(const gprx_problem*, size_t, int32_t, gprx_hip_model** out,
                          gprx_error* err) {
  if (out) *out = nullptr;
  gprx::set_error(err, GPRX_UNSUPPORTED, "batch not built");
  return GPRX_UNSUPPORTED;
}
int gprx_hip_batch_value_and_grad(gprx_hip_model*, const double* const*, double*, double* const*,
                                  gprx_error* errs) {
  gprx::set_error(errs, GPRX_UNSUPPORTED, "batch not built");
  return GPRX_UNSUPPORTED;
}
}
