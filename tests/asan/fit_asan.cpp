// Host-side AddressSanitizer run (see build.sh): batched L-BFGS fits of
// n = 1024, d = 4 points with an RBF kernel and 1 + 16 starts (FP64 without
// and with a jitter policy, mixed), then batches of 48 uneven ARD problems
// in both precisions.  The library's host code is built with
// -fsanitize=address, its device code as usual.
#include <cmath>
#include <cstdio>
#include <cstdint>
#include <vector>
#include "../../include/gprx_hip.h"
int main() {
  const size_t n = 1024, d = 4;
  std::vector<double> x(n * d), y(n);
  uint64_t s = 12345;
  auto u = [&]() { s = s * 6364136223846793005ULL + 1442695040888963407ULL; return (double)(s >> 11) * 0x1.0p-53; };
  for (auto& v : x) v = u();
  for (size_t i = 0; i < n; ++i)
    y[i] = std::sin(3 * x[i]) + std::cos(2 * x[n + i]) + 0.5 * x[2 * n + i] * x[3 * n + i] + 0.1 * (u() - 0.5);
  const int32_t ops[2] = {GPRX_OP_RBF, 0};
  gprx_kernel_program prog{ops, 1};
  double th[2] = {0.0, std::log(0.1)}, best[2];
  gprx_lbfgs_options o;
  gprx_lbfgs_default_options(&o);
  o.restarts = 16;
  double val = 0;
  uint64_t it = 0;
  int32_t bs = -1;
  gprx_error e{};
  gprx_jitter_policy jp{};
  jp.initial = 1e-8;
  jp.multiplier = 10.0;
  jp.max_retries = 5;
  jp.max_jitter = 1e-3;
  for (int rep = 0; rep < 4; ++rep) {
    const int prec = rep == 3 ? GPRX_MIXED : GPRX_FP64;
    const int rc = gprx_hip_exact_fit(x.data(), n, d, y.data(), &prog, th, 2, nullptr, prec, rep >= 1 ? &jp : nullptr, &o,
                                      best, &val, &it, &bs, &e);
    std::printf("fit rc %d value %.12f iterations %llu start %d theta %.6f %.6f\n", rc, val,
                (unsigned long long)it, bs, best[0], best[1]);
    std::fflush(stdout);
  }
  // batched evaluations of uneven problems (ARD), both precisions
  const int K = 48;
  std::vector<std::vector<double>> bx(K), by(K), bth(K), bgr(K);
  std::vector<std::vector<int32_t>> bops(K);
  std::vector<gprx_kernel_program> bprog(K);
  std::vector<gprx_problem> probs(K);
  for (int p = 0; p < K; ++p) {
    const size_t np = 200 + (size_t)(u() * 1900), dp = 1 + (size_t)(u() * 7.99);
    bx[p].resize(np * dp);
    by[p].resize(np);
    for (auto& v : bx[p]) v = u();
    for (auto& v : by[p]) v = u() - 0.5;
    bops[p] = {GPRX_OP_RBF_ARD, (int32_t)dp};
    bprog[p] = gprx_kernel_program{bops[p].data(), 1};
    bth[p].resize(dp + 1);
    for (size_t i = 0; i < dp; ++i) bth[p][i] = std::log(0.2 + u());
    bth[p][dp] = std::log(p % 5 == 0 ? 1e-9 : 1e-2);  // some need FP64 fallbacks in the mixed form
    bgr[p].resize(dp + 1);
    probs[p] = gprx_problem{bx[p].data(), np, dp, by[p].data(), &bprog[p], bth[p].data(), dp + 1};
  }
  for (int prec : {GPRX_FP64, GPRX_MIXED}) {
    gprx_hip_model* m = nullptr;
    int rc = gprx_hip_batch_create(probs.data(), K, prec, &m, &e);
    std::printf("batch create prec %d rc %d\n", prec, rc);
    std::vector<const double*> thp(K);
    std::vector<double*> grp(K);
    std::vector<double> nl(K);
    std::vector<gprx_error> errs(K);
    for (int p = 0; p < K; ++p) { thp[p] = bth[p].data(); grp[p] = bgr[p].data(); }
    for (int r = 0; r < 3; ++r) {
      rc = gprx_hip_batch_value_and_grad(m, thp.data(), nl.data(), grp.data(), errs.data());
      std::printf("batch eval prec %d rc %d nlml0 %.10f nlml5 %.10f\n", prec, rc, nl[0], nl[5]);
      std::fflush(stdout);
    }
    gprx_hip_destroy(m);
  }
  return 0;
}
